// les.hip -- Compute_Smagorinsky_Constant_1 (les.c:75-1141) for gfx950.
//
// les == 1 keeps the constant lCs = cs2 (no kernel: the eddy-viscosity pass
// reads the scalar).  les == 2 runs the dynamic procedure as three passes
// over the DMDA arrays, every FP64 expression in the reference's order:
//
//   k_dyn_grad  per cell: velocity gradient (Compute_du_center +
//               Compute_du_dxyz), |S| and the test-filtered velocity
//               (les.c:206-256) -> 13 scratch planes
//   k_dyn_lm    per cell: the Leonard term L_ij and the model term M_ij from
//               27-point Simpson test filters (rhs2.c:499-523) of the
//               contravariant fluxes times velocity and of S_ij |S|,
//               contracted with the covariant metric tensor into LM and MM
//               (les.c:294-697)
//   k_dyn_cs    per cell: C = 1/2 <LM> / (<MM> + 1e-4) with a second Simpson
//               filter or a homogeneous average (k_dyn_avg_*), clipped to
//               [0, max_cs] (les.c:733-980)
//
// Every pass is one thread per DA point with i fastest, so neighbour reads
// along i coalesce and the 3x3x3 reuse is served by L1/L2; scratch planes
// are SoA.  Periodic ghost cells are not written: readers map them to their
// source cell (the reference copies them explicitly, les.c:262-292,
// 702-731), non-periodic ghosts hold 0 as the reference's VecSet leaves them.
#include <hip/hip_runtime.h>

#include <cmath>
#include <algorithm>
#include <cstdlib>

#include "vfs_internal.h"
#include "vfs_reduce.h"

#pragma clang fp contract(off)

namespace vfs {
namespace {

constexpr int TPB = 256;
constexpr double LES_EPS = 1.e-4;  // les.c:27: this file-scope constant shadows the -les_eps option
constexpr double WALL_CS = 0.001;  // les.c:28
constexpr double SOLID = 0.1;      // k-omega.c:19

// scratch planes (n doubles each)
enum { D_G = 0, D_SABS = 9, D_UF = 10, D_LM = 13, D_MM = 14, D_NPLANES = 15 };
// homogeneous-average modes (les.c:799-958)
enum { H_NONE = 0, H_IK = 1, H_I = 2, H_J = 3, H_K = 4 };

__device__ inline long long pidx(const Grid &g, int k, int j, int i) { return ((long long)k * g.my + j) * g.mx + i; }

// the point whose value the reference copies into a periodic ghost cell
__device__ inline long long src_idx(const Grid &g, int k, int j, int i)
{
    if (g.pi) { if (i == 0) i = g.mx - 2; else if (i == g.mx - 1) i = 1; }
    if (g.pj) { if (j == 0) j = g.my - 2; else if (j == g.my - 1) j = 1; }
    if (g.pk) { if (k == 0) k = g.mz - 2; else if (k == g.mz - 1) k = 1; }
    return pidx(g, k, j, i);
}

__device__ inline bool interior(const Grid &g, int k, int j, int i)
{
    return i >= 1 && i < g.mx - 1 && j >= 1 && j < g.my - 1 && k >= 1 && k < g.mz - 1;
}

// Compute_du_center (k-omega.c:313-431) of an interior cell; u(c, q) returns
// component c of the field at point q (raw DA index of the neighbour)
template <class U>
__device__ inline void du_center(const Grid &g, const double *__restrict__ nv, int k, int j, int i, U u, double d[3][3])
{
    const long long q = pidx(g, k, j, i), P = g.plane, X = g.mx;
    const bool ie = nv[q + 1] > SOLID || (!g.pi && i == g.mx - 2), iw = nv[q - 1] > SOLID || (!g.pi && i == 1);
    const bool jn = nv[q + X] > SOLID || (!g.pj && j == g.my - 2), js = nv[q - X] > SOLID || (!g.pj && j == 1);
    // k-omega.c:390-395 tests k == 0 for the lower one-sided case
    const bool kt = nv[q + P] > SOLID || (!g.pk && k == g.mz - 2), kb = nv[q - P] > SOLID || (!g.pk && k == 0);
    for (int c = 0; c < 3; c++) {
        if (ie) d[c][0] = (u(c, q) - u(c, q - 1));
        else if (iw) d[c][0] = (u(c, q + 1) - u(c, q));
        else d[c][0] = (u(c, q + 1) - u(c, q - 1)) * 0.5;
        if (jn) d[c][1] = (u(c, q) - u(c, q - X));
        else if (js) d[c][1] = (u(c, q + X) - u(c, q));
        else d[c][1] = (u(c, q + X) - u(c, q - X)) * 0.5;
        if (kt) d[c][2] = (u(c, q) - u(c, q - P));
        else if (kb) d[c][2] = (u(c, q + P) - u(c, q));
        else d[c][2] = (u(c, q + P) - u(c, q - P)) * 0.5;
    }
}

// Compute_du_dxyz (k-omega.c:605-618): g[c][r] = d u_c / d x_r
__device__ inline void du_dxyz(const double d[3][3], const double *cs, const double *et, const double *ze, double ajc,
                               double gr[3][3])
{
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) gr[c][r] = (d[c][0] * cs[r] + d[c][1] * et[r] + d[c][2] * ze[r]) * ajc;
}

__device__ inline void strain(const double gr[3][3], double S[3][3])
{
    S[0][0] = 0.5 * (gr[0][0] + gr[0][0]); S[0][1] = 0.5 * (gr[0][1] + gr[1][0]); S[0][2] = 0.5 * (gr[0][2] + gr[2][0]);
    S[1][0] = S[0][1]; S[1][1] = 0.5 * (gr[1][1] + gr[1][1]); S[1][2] = 0.5 * (gr[1][2] + gr[2][1]);
    S[2][0] = S[0][2]; S[2][1] = S[1][2]; S[2][2] = 0.5 * (gr[2][2] + gr[2][2]);
}

__device__ inline double strain_abs(const double S[3][3])
{
    return sqrt(2.0 * (S[0][0] * S[0][0] + S[0][1] * S[0][1] + S[0][2] * S[0][2] + S[1][0] * S[1][0] + S[1][1] * S[1][1] +
                       S[1][2] * S[1][2] + S[2][0] * S[2][0] + S[2][1] * S[2][1] + S[2][2] * S[2][2]));
}

// Simpson weight of offset o in {-1, 0, 1} along one direction
__device__ inline double simpson_factor(double sw, int o) { return o == 0 ? sw * 4. : sw; }

// pass 1 (les.c:206-256)
__global__ void __launch_bounds__(TPB) k_dyn_grad(Grid g, const double *__restrict__ ucat, const double *__restrict__ csi,
                                                  const double *__restrict__ eta, const double *__restrict__ zet,
                                                  const double *__restrict__ aj, const double *__restrict__ nvert,
                                                  double *__restrict__ D)
{
    const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (q >= g.n) return;
    const long long n = g.n;
    const int t32 = (int)q, i = t32 % g.mx, j = (t32 / g.mx) % g.my, k = t32 / (g.mx * g.my);
    if (!interior(g, k, j, i) || nvert[q] > 1.1) {
        for (int p = 0; p < D_NPLANES - 2; p++) D[p * n + q] = 0;
        return;
    }
    double d[3][3], gr[3][3], S[3][3];
    du_center(g, nvert, k, j, i, [&](int c, long long s) { return ucat[3 * s + c]; }, d);
    du_dxyz(d, csi + 3 * q, eta + 3 * q, zet + 3 * q, aj[q], gr);
    strain(gr, S);
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) D[(D_G + 3 * c + r) * n + q] = gr[c][r];
    D[D_SABS * n + q] = strain_abs(S);
    // integrate_testfilter_simpson of u, v, w with get_weight(nv = 0.1) weights
    double wsum = 0, vs[3] = {0, 0, 0};
    for (int r = -1; r <= 1; r++)
        for (int qq = -1; qq <= 1; qq++)
            for (int pp = -1; pp <= 1; pp++) {
                const long long s = q + r * g.plane + qq * g.mx + pp;
                const double w = nvert[s] > 0.1 ? 0. : 1. / aj[s];
                const double sw = simpson_factor(simpson_factor(simpson_factor(1.0, r), qq), pp);
                wsum += sw * w;
                for (int c = 0; c < 3; c++) vs[c] += sw * w * ucat[3 * s + c];
            }
    for (int c = 0; c < 3; c++) D[(D_UF + c) * n + q] = vs[c] / wsum;
}

// pass 2 (les.c:294-697): LM and MM of every cell
__global__ void __launch_bounds__(TPB) k_dyn_lm(Grid g, const double *__restrict__ ucat, const double *__restrict__ csi,
                                                const double *__restrict__ eta, const double *__restrict__ zet,
                                                const double *__restrict__ aj, const double *__restrict__ nvert,
                                                double *__restrict__ D)
{
    const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (q >= g.n) return;
    const long long n = g.n;
    const int t32 = (int)q, i = t32 % g.mx, j = (t32 / g.mx) % g.my, k = t32 / (g.mx * g.my);
    if (!interior(g, k, j, i) || nvert[q] > 1.1) {
        D[D_LM * n + q] = 0;
        D[D_MM * n + q] = 0;
        return;
    }
    // get_weight (les.c:30-73): 1/aj of each of the 27 cells, 0 where nvert > 0.1
    auto weight = [&](long long s) { return nvert[s] > 0.1 ? 0. : 1. / aj[s]; };
    // les.c:428-446: test-filter width from the weights (i offset outermost)
    double sum_weight = 0;
    for (int pp = -1; pp <= 1; pp++)
        for (int qq = -1; qq <= 1; qq++)
            for (int r = -1; r <= 1; r++) {
                const double cf = (r == 0 ? 1.0 : 0.5) * (qq == 0 ? 1.0 : 0.5) * (pp == 0 ? 1.0 : 0.5);
                sum_weight += weight(q + r * g.plane + qq * g.mx + pp) * cf;
            }
    // Simpson filters of U_a u_b and S_ab |S| over the 27 cells (k offset outermost)
    double wsum = 0, fUu[9], fSS[9];
    for (int a = 0; a < 9; a++) fUu[a] = fSS[a] = 0;
    for (int r = -1; r <= 1; r++)
        for (int qq = -1; qq <= 1; qq++)
            for (int pp = -1; pp <= 1; pp++) {
                const long long s = q + r * g.plane + qq * g.mx + pp;
                const long long sr = src_idx(g, k + r, j + qq, i + pp);
                const double uc[3] = {ucat[3 * s], ucat[3 * s + 1], ucat[3 * s + 2]};
                const double *cs = csi + 3 * s, *et = eta + 3 * s, *ze = zet + 3 * s;
                double U[3];
                U[0] = uc[0] * cs[0] + uc[1] * cs[1] + uc[2] * cs[2];
                U[1] = uc[0] * et[0] + uc[1] * et[1] + uc[2] * et[2];
                U[2] = uc[0] * ze[0] + uc[1] * ze[1] + uc[2] * ze[2];
                double gr[3][3], S[3][3];
                for (int c = 0; c < 3; c++)
                    for (int rr = 0; rr < 3; rr++) gr[c][rr] = D[(D_G + 3 * c + rr) * n + sr];
                strain(gr, S);
                const double sa = D[D_SABS * n + sr];
                const double sww = simpson_factor(simpson_factor(simpson_factor(1.0, r), qq), pp) * weight(s);
                wsum += sww;
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) {
                        fUu[3 * a + b] += sww * (U[a] * uc[b]);
                        fSS[3 * a + b] += sww * (S[a][b] * sa);
                    }
            }
    const double filter = pow(1. / aj[q], 1. / 3.);
    const double test_filter = pow(sum_weight, 1. / 3.);
    const double *c0 = csi + 3 * q, *e0 = eta + 3 * q, *z0 = zet + 3 * q;
    const double fu[3] = {D[D_UF * n + q], D[(D_UF + 1) * n + q], D[(D_UF + 2) * n + q]};
    double fU[3];
    fU[0] = fu[0] * c0[0] + fu[1] * c0[1] + fu[2] * c0[2];
    fU[1] = fu[0] * e0[0] + fu[1] * e0[1] + fu[2] * e0[2];
    fU[2] = fu[0] * z0[0] + fu[1] * z0[1] + fu[2] * z0[2];
    // gradient of the filtered velocity at the cell (les.c:475-486)
    double d[3][3], gh[3][3], Sh[3][3];
    du_center(g, nvert, k, j, i,
              [&](int c, long long s) {
                  const int t = (int)s, ii = t % g.mx, jj = (t / g.mx) % g.my, kk = t / (g.mx * g.my);
                  return D[(D_UF + c) * n + src_idx(g, kk, jj, ii)];
              },
              d);
    du_dxyz(d, c0, e0, z0, aj[q], gh);
    strain(gh, Sh);
    const double S_hat = strain_abs(Sh);
    double L[3][3], Mc[3][3], M[3][3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) L[a][b] = fUu[3 * a + b] / wsum - fU[a] * fu[b];
    const double gm[3][3] = {{c0[0], c0[1], c0[2]}, {e0[0], e0[1], e0[2]}, {z0[0], z0[1], z0[2]}};
    double gc[3][3], G[3][3];
    covariant(gm, gc);
    G[0][0] = gc[0][0] * gc[0][0] + gc[1][0] * gc[1][0] + gc[2][0] * gc[2][0];
    G[1][1] = gc[0][1] * gc[0][1] + gc[1][1] * gc[1][1] + gc[2][1] * gc[2][1];
    G[2][2] = gc[0][2] * gc[0][2] + gc[1][2] * gc[1][2] + gc[2][2] * gc[2][2];
    G[0][1] = G[1][0] = gc[0][1] * gc[0][0] + gc[1][1] * gc[1][0] + gc[2][1] * gc[2][0];
    G[0][2] = G[2][0] = gc[0][2] * gc[0][0] + gc[1][2] * gc[1][0] + gc[2][2] * gc[2][0];
    G[1][2] = G[2][1] = gc[0][1] * gc[0][2] + gc[1][1] * gc[1][2] + gc[2][1] * gc[2][2];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            Mc[a][b] = -pow(test_filter, 2.) * S_hat * Sh[a][b] + pow(filter, 2.) * (fSS[3 * a + b] / wsum);
    for (int a = 0; a < 3; a++) {
        M[a][0] = Mc[a][0] * c0[0] + Mc[a][1] * c0[1] + Mc[a][2] * c0[2];
        M[a][1] = Mc[a][0] * e0[0] + Mc[a][1] * e0[1] + Mc[a][2] * e0[2];
        M[a][2] = Mc[a][0] * z0[0] + Mc[a][1] * z0[1] + Mc[a][2] * z0[2];
    }
    double num = 0, denom = 0;
    for (int qq = 0; qq < 3; qq++)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) num += L[b][a] * M[a][qq] * G[b][qq];
    for (int m = 0; m < 3; m++)
        for (int nn = 0; nn < 3; nn++)
            for (int l = 0; l < 3; l++) denom += M[nn][m] * M[nn][l] * G[m][l];
    D[D_LM * n + q] = num;
    D[D_MM * n + q] = denom;
}

// i and k homogeneous (les.c:799-838): block j averages LM and MM over the
// fluid cells of plane j
__global__ void __launch_bounds__(RED_THREADS) k_dyn_avg_plane(Grid g, const double *__restrict__ nvert,
                                                               const double *__restrict__ D, double *__restrict__ avg)
{
    const int j = blockIdx.x;
    const long long n = g.n;
    double lm = 0, mm = 0, cnt = 0;
    if (j >= 1 && j < g.my - 1) {
        const int ni = g.mx - 2;
        const long long tot = (long long)(g.mz - 2) * ni;
        for (long long t = threadIdx.x; t < tot; t += blockDim.x) {
            const int k = 1 + (int)(t / ni), i = 1 + (int)(t % ni);
            const long long q = pidx(g, k, j, i);
            if (nvert[q] < 0.1) { lm += D[D_LM * n + q]; mm += D[D_MM * n + q]; cnt += 1; }
        }
    }
    lm = block_reduce<OpSum>(lm);
    mm = block_reduce<OpSum>(mm);
    cnt = block_reduce<OpSum>(cnt);
    if (threadIdx.x == 0) {
        if (cnt > 0) { lm /= cnt; mm /= cnt; }
        avg[2 * j] = lm;
        avg[2 * j + 1] = mm;
    }
}

// one homogeneous direction (les.c:840-955): one wavefront per line of
// interior cells along `dir`; lines numbered (a-1)*(nb-2) + (b-1) over the
// two other directions (a = k, b = j for i lines; a = k, b = i for j lines;
// a = j, b = i for k lines)
__device__ inline void line_dims(const Grid &g, int dir, int &len, int &na, int &nb)
{
    len = dir == 0 ? g.mx : dir == 1 ? g.my : g.mz;
    na = dir == 2 ? g.my : g.mz;
    nb = dir == 0 ? g.my : g.mx;
}

__device__ inline long long line_point(const Grid &g, int dir, int a, int b, int t)
{
    const int k = dir == 2 ? t : a, j = dir == 0 ? b : dir == 1 ? t : a, i = dir == 0 ? t : b;
    return pidx(g, k, j, i);
}

__global__ void __launch_bounds__(256) k_dyn_avg_line(Grid g, int dir, const double *__restrict__ nvert,
                                                      const double *__restrict__ D, double *__restrict__ avg)
{
    int len, na, nb;
    line_dims(g, dir, len, na, nb);
    const long long line = (blockIdx.x * (long long)blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const long long nlines = (long long)(na - 2) * (nb - 2);
    if (line >= nlines) return;  // whole wavefronts leave together
    const int a = 1 + (int)(line / (nb - 2)), b = 1 + (int)(line % (nb - 2));
    const long long n = g.n;
    double lm = 0, mm = 0, cnt = 0;
    for (int t = 1 + lane; t < len - 1; t += 64) {
        const long long q = line_point(g, dir, a, b, t);
        if (nvert[q] < 0.1) { lm += D[D_LM * n + q]; mm += D[D_MM * n + q]; cnt += 1; }
    }
    lm = wave_reduce<OpSum>(lm);
    mm = wave_reduce<OpSum>(mm);
    cnt = wave_reduce<OpSum>(cnt);
    if (lane == 0) {
        if (cnt > 0) { lm /= cnt; mm /= cnt; }
        avg[2 * line] = lm;
        avg[2 * line + 1] = mm;
    }
}

// pass 3 (les.c:733-797, 960-980): Cs of every point; periodic ghosts are
// filled afterwards by the periodic fix (les.c:1092-1118)
__global__ void __launch_bounds__(TPB) k_dyn_cs(Grid g, int homo, double max_cs, const double *__restrict__ aj,
                                                const double *__restrict__ nvert, const double *__restrict__ D,
                                                const double *__restrict__ avg, double *__restrict__ cs)
{
    const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (q >= g.n) return;
    const long long n = g.n;
    const int t32 = (int)q, i = t32 % g.mx, j = (t32 / g.mx) % g.my, k = t32 / (g.mx * g.my);
    const double nvq = nvert[q];
    if (!interior(g, k, j, i) || nvq > 1.1) {
        cs[q] = 0;
        return;
    }
    double c;
    if (homo == H_NONE) {
        double wsum = 0, vl = 0, vm = 0;
        for (int r = -1; r <= 1; r++)
            for (int qq = -1; qq <= 1; qq++)
                for (int pp = -1; pp <= 1; pp++) {
                    int K = k + r, J = j + qq, I = i + pp;
                    const long long s = pidx(g, K, J, I);
                    double w = 1. / aj[s];
                    if (nvert[s] > 1.1) w = 0;
                    if (g.pi) { if (I == 0) I = g.mx - 2; else if (I == g.mx - 1) I = 1; }
                    else if (I == 0 || I == g.mx - 1) w = 0;
                    if (g.pj) { if (J == 0) J = g.my - 2; else if (J == g.my - 1) J = 1; }
                    else if (J == 0 || j == g.my - 1) w = 0;  // les.c:763 tests the centre's j
                    if (g.pk) { if (K == 0) K = g.mz - 2; else if (K == g.mz - 1) K = 1; }
                    else if (K == 0 || K == g.mz - 1) w = 0;
                    const long long sr = pidx(g, K, J, I);
                    const double sw = simpson_factor(simpson_factor(simpson_factor(1.0, r), qq), pp);
                    wsum += sw * w;
                    vl += sw * w * D[D_LM * n + sr];
                    vm += sw * w * D[D_MM * n + sr];
                }
        const double C = 0.5 * (vl / wsum) / ((vm / wsum) + LES_EPS);
        c = C < 0 ? 0 : C;  // PetscMax(C, 0)
    } else {
        const double C = 0.5 * D[D_LM * n + q] / (D[D_MM * n + q] + LES_EPS);
        c = C < 0 ? 0 : C;
        if (homo == H_IK) {
            c = 0.5 * avg[2 * j] / (avg[2 * j + 1] + LES_EPS);
        } else if (nvq < 0.1) {
            const int dir = homo - H_I;
            int len, na, nb;
            line_dims(g, dir, len, na, nb);
            const int a = dir == 2 ? j : k, b = dir == 0 ? j : i;
            const long long line = (long long)(a - 1) * (nb - 2) + (b - 1);
            c = 0.5 * avg[2 * line] / (avg[2 * line + 1] + LES_EPS);
        }
    }
    // PetscMax(a, b) = a < b ? b : a; PetscMin(a, b) = a < b ? a : b
    if (nvq > 0.1 && nvq < 1.1) c = WALL_CS < c ? c : WALL_CS;
    c = c < 0 ? 0 : c;
    cs[q] = c < max_cs ? c : max_cs;
}

__global__ void k_fill(double *__restrict__ f, long long n, double v)
{
    const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (q < n) f[q] = v;
}

int homo_mode(const vfs_params &p)
{
    if (p.i_homo_filter && p.k_homo_filter) return H_IK;
    if (p.i_homo_filter) return H_I;
    if (p.j_homo_filter) return H_J;
    if (p.k_homo_filter) return H_K;
    return H_NONE;
}

}  // namespace

int dynamic_smagorinsky(vfs_ctx *c)
{
    const Grid &g = c->g;
    const long long n = g.n;
    const long long pmax = std::max((long long)g.my * g.mz, std::max((long long)g.mx * g.mz, (long long)g.mx * g.my));
    if (!c->dyn) {
        const size_t bytes = ((size_t)D_NPLANES * n + 2 * (size_t)pmax) * sizeof(double);
        hipError_t e = hipMalloc((void **)&c->dyn, bytes);
        if (e != hipSuccess) { c->dyn = nullptr; return check_hip(e, "hipMalloc (dynamic Smagorinsky scratch)"); }
    }
    double *D = c->dyn, *avg = c->dyn + (size_t)D_NPLANES * n;
    hipStream_t st = c->stream;
    const unsigned nb = blocks_for(n, TPB);
    hipLaunchKernelGGL(k_dyn_grad, dim3(nb), dim3(TPB), 0, st, g, c->f[VFS_UCAT], c->f[VFS_CSI], c->f[VFS_ETA],
                       c->f[VFS_ZET], c->f[VFS_AJ], c->f[VFS_NVERT], D);
    hipLaunchKernelGGL(k_dyn_lm, dim3(nb), dim3(TPB), 0, st, g, c->f[VFS_UCAT], c->f[VFS_CSI], c->f[VFS_ETA],
                       c->f[VFS_ZET], c->f[VFS_AJ], c->f[VFS_NVERT], D);
    VFS_LAUNCH_CHECK("dynamic Smagorinsky LM/MM");
    const int homo = homo_mode(c->p);
    if (homo == H_IK) {
        hipLaunchKernelGGL(k_dyn_avg_plane, dim3(g.my), dim3(RED_THREADS), 0, st, g, c->f[VFS_NVERT], D, avg);
    } else if (homo != H_NONE) {
        const int dir = homo - H_I;
        const int na = dir == 2 ? g.my : g.mz, nbb = dir == 0 ? g.my : g.mx;
        const long long lines = (long long)(na - 2) * (nbb - 2);
        hipLaunchKernelGGL(k_dyn_avg_line, dim3(blocks_for(lines * 64, 256)), dim3(256), 0, st, g, dir, c->f[VFS_NVERT], D,
                           avg);
    }
    hipLaunchKernelGGL(k_dyn_cs, dim3(nb), dim3(TPB), 0, st, g, homo, c->p.max_cs, c->f[VFS_AJ], c->f[VFS_NVERT], D, avg,
                       c->f[VFS_CS]);
    VFS_LAUNCH_CHECK("dynamic Smagorinsky Cs");
    return periodic_fix(c, c->f[VFS_CS], 1, 0);
}

int fill_field(vfs_ctx *c, double *f, long long n, double v)
{
    hipLaunchKernelGGL(k_fill, dim3(blocks_for(n, TPB)), dim3(TPB), 0, c->stream, f, n, v);
    VFS_LAUNCH_CHECK("fill");
    return VFS_OK;
}

// Compute_Smagorinsky_Constant_1 (les.c:75-1141)
int smagorinsky_constant(vfs_ctx *c, int ti, int tistart, int rstart)
{
    if (ti < 2 && tistart == 0 && !rstart) {  // les.c:77-80
        c->cs_mode = CS_ZERO;
        return VFS_OK;
    }
    if (c->p.les != 2) {  // les.c:82-85: lCs = 0.01 (cs2) everywhere
        c->cs_mode = CS_PARAM;
        return VFS_OK;
    }
    int rc = dynamic_smagorinsky(c);
    if (rc) return rc;
    c->cs_mode = CS_FIELD;
    return VFS_OK;
}

// VFS_CS as the eddy-viscosity pass will read it
int materialize_cs(vfs_ctx *c)
{
    if (c->cs_mode == CS_FIELD) return VFS_OK;
    int rc = fill_field(c, c->f[VFS_CS], c->g.n, c->cs_mode == CS_ZERO ? 0.0 : c->p.cs2);
    if (rc) return rc;
    c->cs_mode = CS_FIELD;
    return VFS_OK;
}

}  // namespace vfs
