// alm.hip -- actuator-line turbine model (rotor_model.c) for gfx950.
//
// One turbine per vfs_alm, line elements between consecutive blade nodes.
// The blade kinematics (rotation about the rotor axis, element centres and
// lengths, twist/chord by radius) touch a few hundred numbers and run on the
// host in FP64; everything that reads or writes the Eulerian grid runs on the
// device:
//   k_alm_nearest   Pre_process      nearest grid node along the first
//                                    interior grid lines, +-6 cell box
//   k_alm_ulagr     Calc_U_lagr      one block per element, dfunc_s3h-weighted
//                                    sum of ucat over its box (tree reduction)
//   k_alm_flagr     Calc_F_lagr_ACL  one thread per element
//   k_alm_feul      Calc_F_eul       one thread per cell of the union of the
//                                    element boxes, gathering every element
//                                    whose box holds the cell in element
//                                    order, so each cell sums its terms in
//                                    the reference's order; then the
//                                    blanking / boundary-layer zeroing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "vfs_internal.h"

#pragma clang fp contract(off)

namespace vfs {

constexpr int ALM_BOX = 6;  // Itpwidth = (int)halfwidth_dfunc + 2 with its default 4.0, rotor_model.c:47
constexpr int ALM_TPB = 256;

// ---- delta kernels (rotor_model.c:5117-5234), truncated pi as the reference
__host__ __device__ inline double dfunc_2h(double r) { return fabs(r) < 1.0 ? 1.0 - fabs(r) : 0.0; }

__host__ __device__ inline double dfunc_s3h(double r)
{
    const double a = fabs(r);
    if (a <= 1.0)
        return 17.0 / 48.0 + sqrt(3.0) * 3.14159265 / 108.0 + a / 4.0 - r * r / 4.0 +
               (1.0 - 2.0 * a) * sqrt(-12.0 * r * r + 12.0 * a + 1.0) / 16.0 -
               sqrt(3.0) * asin(sqrt(3.0) * (2.0 * a - 1.0) / 2.0) / 12.0;
    if (a >= 1.0 && a <= 2.0)
        return 55.0 / 48.0 - sqrt(3.0) * 3.14159265 / 108.0 - 13.0 * a / 12.0 + r * r / 4.0 +
               (2.0 * a - 3.0) * sqrt(-12.0 * r * r + 36.0 * a - 23.0) / 48.0 +
               sqrt(3.0) * asin(sqrt(3.0) * (2.0 * a - 3.0) / 2.0) / 36.0;
    return 0.0;
}

__host__ __device__ inline double dfunc_s4h(double r)
{
    const double a = fabs(r);
    if (a <= 0.5) return 3.0 / 8.0 + 3.14159265 / 32.0 - r * r / 4.0;
    if (a >= 0.5 && a <= 1.5)
        return 1.0 / 4.0 + (1.0 - a) * sqrt(-2.0 + 8.0 * a - 4.0 * (r * r)) / 8.0 - asin(sqrt(2.0) * (a - 1.0)) / 8.0;
    if (a >= 1.5 && a <= 2.5)
        return 17.0 / 16.0 - 3.14159265 / 64.0 - 3.0 * a / 4.0 + r * r / 8.0 +
               (a - 2.0) * sqrt(-14.0 + 16.0 * a - 4.0 * (r * r)) / 16.0 + asin(sqrt(2.0) * (a - 2.0)) / 16.0;
    return 0.0;
}

__host__ __device__ inline double dfunc_exp(double r, double n)
{
    return exp(-(r / n) * (r / n)) / (n * sqrt(3.1415926));
}

// spacing of cell (k,j,i) along its three grid directions, rotor_model.c:1935-1940
__device__ inline void cell_spacing(const double *csi, const double *eta, const double *zet, const double *aj, long long q,
                                    double &hx, double &hy, double &hz)
{
    const double *c = csi + 3 * q, *e = eta + 3 * q, *z = zet + 3 * q;
    hx = 1.0 / aj[q] / sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    hy = 1.0 / aj[q] / sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    hz = 1.0 / aj[q] / sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
}

// unit covariant base vectors (Calculate_normal, rhs2.c:614-647)
__device__ inline void cell_normals(const double *csi, const double *eta, const double *zet, long long q, double n[3][3])
{
    double gm[3][3], G[3][3];
    for (int c = 0; c < 3; c++) { gm[0][c] = csi[3 * q + c]; gm[1][c] = eta[3 * q + c]; gm[2][c] = zet[3 * q + c]; }
    covariant(gm, G);
    for (int d = 0; d < 3; d++) {
        const double x = G[0][d], y = G[1][d], z = G[2][d];
        const double s = sqrt(x * x + y * y + z * z);
        n[d][0] = x / s; n[d][1] = y / s; n[d][2] = z / s;
    }
}

struct GridMet {
    Grid g;
    const double *coor, *csi, *eta, *zet, *aj, *nvert;
};

__device__ inline double CO(const GridMet &m, int c, int k, int j, int i) { return m.coor[3 * gidx(m.g, k, j, i) + c]; }

// Pre_process, rotor_model.c:1880-1995 (one rank)
__global__ void k_alm_nearest(GridMet m, int n, const double *__restrict__ cent, int *__restrict__ box)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const Grid &g = m.g;
    const int lo[3] = {1, 1, 1}, hi[3] = {g.mx - 1, g.my - 1, g.mz - 1};
    const double cx = cent[3 * l], cy = cent[3 * l + 1], cz = cent[3 * l + 2];
    int cc[3] = {0, 0, 0};
    for (int d = 0; d < 3; d++) {
        double dmin = 1.e6;
        for (int s = lo[d]; s < hi[d]; s++) {
            const int i = d == 0 ? s : lo[0] + 1, j = d == 1 ? s : lo[1] + 1, k = d == 2 ? s : lo[2] + 1;
            const double r1 = cx - CO(m, 0, k, j, i), r2 = cy - CO(m, 1, k, j, i), r3 = cz - CO(m, 2, k, j, i);
            const double d1 = sqrt(r1 * r1 + r2 * r2 + r3 * r3);
            if (d1 < dmin) { dmin = d1; cc[d] = s; }
        }
    }
    const double r1 = cx - CO(m, 0, cc[2], cc[1], cc[0]), r2 = cy - CO(m, 1, cc[2], cc[1], cc[0]),
                 r3 = cz - CO(m, 2, cc[2], cc[1], cc[0]);
    const double dmin = sqrt(r1 * r1 + r2 * r2 + r3 * r3);
    double hx, hy, hz;
    cell_spacing(m.csi, m.eta, m.zet, m.aj, gidx(g, cc[2], cc[1], cc[0]), hx, hy, hz);
    const double dh_grid = 2.0 * sqrt(hx * hx + hy * hy + hz * hz);
    if (dmin - dh_grid > 1.e-6) cc[0] = cc[1] = cc[2] = 0;  // the single-rank reference divides by a zero count here
    for (int d = 0; d < 3; d++) {
        const int c1 = cc[d] - ALM_BOX, c2 = cc[d] + ALM_BOX;
        if (c1 > hi[d] || c2 < lo[d]) { box[6 * l + 2 * d] = lo[d]; box[6 * l + 2 * d + 1] = lo[d]; }
        else {
            box[6 * l + 2 * d] = c1 > lo[d] ? c1 : lo[d];
            box[6 * l + 2 * d + 1] = c2 < hi[d] ? c2 : hi[d];
        }
    }
}

// Calc_U_lagr, rotor_model.c:2977-3036: one block per element
__global__ void __launch_bounds__(ALM_TPB) k_alm_ulagr(GridMet m, const double *__restrict__ ucat, int n,
                                                       const double *__restrict__ cent, const int *__restrict__ box,
                                                       double *__restrict__ ulagr)
{
    const int l = blockIdx.x;
    if (l >= n) return;
    const int *b = box + 6 * l;
    const int ni = b[1] - b[0], nj = b[3] - b[2], nk = b[5] - b[4];
    const int cells = ni > 0 && nj > 0 && nk > 0 ? ni * nj * nk : 0;
    double u[3] = {0, 0, 0};
    for (int t = threadIdx.x; t < cells; t += blockDim.x) {
        const int i = b[0] + t % ni, j = b[2] + (t / ni) % nj, k = b[4] + t / (ni * nj);
        double xc[3];
        for (int c = 0; c < 3; c++)
            xc[c] = 0.125 * (CO(m, c, k, j, i) + CO(m, c, k, j - 1, i) + CO(m, c, k - 1, j, i) + CO(m, c, k - 1, j - 1, i) +
                             CO(m, c, k, j, i - 1) + CO(m, c, k, j - 1, i - 1) + CO(m, c, k - 1, j, i - 1) +
                             CO(m, c, k - 1, j - 1, i - 1));
        const long long q = gidx(m.g, k, j, i);
        double hx, hy, hz, nn[3][3];
        cell_spacing(m.csi, m.eta, m.zet, m.aj, q, hx, hy, hz);
        const double rx = xc[0] - cent[3 * l], ry = xc[1] - cent[3 * l + 1], rz = xc[2] - cent[3 * l + 2];
        cell_normals(m.csi, m.eta, m.zet, q, nn);
        const double r1 = (rx * nn[0][0] + ry * nn[0][1] + rz * nn[0][2]) / hx;
        const double r2 = (rx * nn[1][0] + ry * nn[1][1] + rz * nn[1][2]) / hy;
        const double r3 = (rx * nn[2][0] + ry * nn[2][1] + rz * nn[2][2]) / hz;
        const double df = dfunc_s3h(r1) * dfunc_s3h(r2) * dfunc_s3h(r3);
        for (int c = 0; c < 3; c++) u[c] += ucat[3 * q + c] * df;
    }
    __shared__ double red[3][ALM_TPB];
    for (int c = 0; c < 3; c++) red[c][threadIdx.x] = u[c];
    __syncthreads();
    for (int s = ALM_TPB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int c = 0; c < 3; c++) red[c][threadIdx.x] += red[c][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; c++) ulagr[3 * l + c] = red[c][0];
}

// polar tables of all foils, flattened
struct Polars {
    int nfoil;
    const int *off;     // per foil: n_ac, off_ac, n_cl, off_cl, n_cd, off_cd (6 ints)
    const double *tab;  // r_ac | twist | chord for ac; ang | val for cl / cd
};

__device__ inline double polar(double aoa, int nq, const double *ang, const double *val, double lim, bool lift, double v)
{
    const double pi = 3.141592653589793;
    bool inrange = false;
    for (int j = 0; j < nq - 1; j++)
        if (aoa >= ang[j] && aoa <= ang[j + 1]) {
            const double fac1 = (ang[j + 1] - aoa) / (ang[j + 1] - ang[j]);
            const double fac2 = (-ang[j] + aoa) / (ang[j + 1] - ang[j]);
            v = fac1 * val[j] + fac2 * val[j + 1];
            inrange = true;
        }
    if (!inrange) {
        if (aoa < ang[0] && aoa > -45.0) v = (-lim - val[0]) * (aoa - ang[0]) / (-45.0 - ang[0]) + val[0];
        if (aoa <= -45.0 && aoa >= -90.0) v = lift ? lim * sin(2.0 * aoa * pi / 180.0) : -lim;
        if (aoa < -90.0) v = lift ? 0.0 : -lim;
        if (aoa > ang[nq - 1] && aoa < 45.0) v = (lim - val[nq - 1]) * (aoa - ang[nq - 1]) / (45.0 - ang[nq - 1]) + val[nq - 1];
        if (aoa >= 45.0 && aoa <= 90.0) v = lift ? lim * sin(2.0 * aoa * pi / 180.0) : lim;
    }
    return v;
}

struct TurbineDev {
    double nx, ny, nz, xc, yc, zc, pitch, r_nacelle;
    int nb, nvb;
    // optional corrections (rotor_model.c:2488-2558)
    int shen, prandtl, shen1, c3d, num_blade, nacelle_here;
    double r_rotor, tsr, a_shen, b_shen, c_shen, k_shen1, c0_cl;
};

// tip-loss factor F1 of the selected models, rotor_model.c:2488-2544: the later
// model wins; these branches use pi = 3.1415926
__device__ inline double tip_loss(const TurbineDev &t, double r, double aoa, double twist)
{
    double F1 = 1.0;
    if (t.shen) {
        const double R = t.r_rotor, pi1 = 3.1415926;
        const double phi1 = (aoa + twist + t.pitch) * pi1 / 180.0;
        const double g1 = exp(-t.a_shen * (t.num_blade * fabs(t.tsr) - t.b_shen)) + t.c_shen;
        F1 = 2.0 * acos(exp(-g1 * t.num_blade * (fmax(R - r, 0.0)) / 2.0 / r / fabs(sin(phi1)))) / pi1;
    }
    if (t.prandtl) {
        const double R = t.r_rotor, pi1 = 3.1415926;
        const double g1 = sqrt(1 + pow(t.tsr, 2));
        F1 = 2.0 * acos(exp(-g1 * t.num_blade * (fmax(R - r, 0.0)) / 2.0 / R)) / pi1;
    }
    if (t.shen1) {
        const double R = t.r_rotor, pi1 = 3.1415926;
        const double phi1 = (aoa + twist + t.pitch) * pi1 / 180.0;
        const double lam = fabs(t.tsr);
        double g1;
        if (lam < 7.0636) g1 = -1.7636 * lam + 11.6112;
        else if (lam > 7.0636 && lam < 9.8146) g1 = -0.5805 * lam + 3.2542;
        else g1 = -0.5077 * lam + 2.5397;
        g1 *= t.k_shen1;
        g1 = exp(g1);
        F1 = 2.0 * acos(exp(-g1 * t.num_blade * (fmax(R - r, 0.0)) / 2.0 / r / fabs(sin(phi1)))) / pi1;
    }
    return F1;
}

// Calc_F_lagr_ACL, rotor_model.c:2384-2597 (rotor_model 3/6): blade-element
// forces with the selected tip-loss factor and 3D lift correction; the
// nacelle zeroing here unless the force smoothing runs first
__global__ void k_alm_flagr(TurbineDev t, Polars P, const double *__restrict__ cent, const double *__restrict__ ulagr,
                            const double *__restrict__ unode, const double *__restrict__ twist,
                            const double *__restrict__ chord, double *__restrict__ aoa, double *__restrict__ cl,
                            double *__restrict__ cd, double *__restrict__ flagr)
{
    const int ne = t.nvb - 1;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= t.nb * ne) return;
    const double pi = 3.141592653589793;
    const int n1 = (l / ne) * t.nvb + l % ne, n2 = n1 + 1;
    const double rx = cent[3 * l] - t.xc, ry = cent[3 * l + 1] - t.yc, rz = cent[3 * l + 2] - t.zc;
    const double r = sqrt(rx * rx + ry * ry + rz * rz) + 1.e-20;
    const double nbl[3] = {rx / r, ry / r, rz / r};
    const double ux = 0.5 * (unode[3 * n1] + unode[3 * n2]);
    const double uy = 0.5 * (unode[3 * n1 + 1] + unode[3 * n2 + 1]);
    const double uz = 0.5 * (unode[3 * n1 + 2] + unode[3 * n2 + 2]);
    const double ublade = sqrt(ux * ux + uy * uy + uz * uz);
    const double nrot[3] = {ux / (ublade + 1.e-20), uy / (ublade + 1.e-20), uz / (ublade + 1.e-20)};
    double ub[3] = {ulagr[3 * l] - ux, ulagr[3 * l + 1] - uy, ulagr[3 * l + 2] - uz};
    const double ur = ub[0] * nbl[0] + ub[1] * nbl[1] + ub[2] * nbl[2];
    ub[0] -= ur * nbl[0];
    ub[1] -= ur * nbl[1];
    ub[2] -= ur * nbl[2];
    const double uref = sqrt(ub[0] * ub[0] + ub[1] * ub[1] + ub[2] * ub[2]);
    const double nrel[3] = {ub[0] / (uref + 1.e-20), ub[1] / (uref + 1.e-20), ub[2] / (uref + 1.e-20)};
    double tmp = -nrel[0] * nrot[0] - nrel[1] * nrot[1] - nrel[2] * nrot[2];
    tmp /= (1 + 1e-9);
    const double a = acos(tmp) * 180.0 / pi - twist[l] - t.pitch;
    aoa[l] = a;
    double nL[3] = {nrel[1] * nbl[2] - nrel[2] * nbl[1], nrel[2] * nbl[0] - nrel[0] * nbl[2], nrel[0] * nbl[1] - nrel[1] * nbl[0]};
    tmp = t.nx * nL[0] + t.ny * nL[1] + t.nz * nL[2];
    if (tmp < 0.0) { nL[0] = -nL[0]; nL[1] = -nL[1]; nL[2] = -nL[2]; }
    double vcl = cl[l], vcd = cd[l];
    for (int f = 0; f < P.nfoil; f++) {
        const int *o = P.off + 6 * f;
        const double *rac = P.tab + o[1];
        if (r >= rac[0] && r <= rac[o[0] - 1]) {
            vcd = polar(a, o[4], P.tab + o[5], P.tab + o[5] + o[4], 1.2, false, vcd);
            vcl = polar(a, o[2], P.tab + o[3], P.tab + o[3] + o[2], 1.05, true, vcl);
        }
    }
    cl[l] = vcl;
    cd[l] = vcd;
    const double factor = chord[l];
    double F[3];
    for (int c = 0; c < 3; c++) F[c] = -0.5 * uref * uref * fabs(vcd) * nrel[c] * factor;
    for (int c = 0; c < 3; c++) F[c] += -0.5 * uref * uref * fabs(vcl) * nL[c] * factor;
    const double F1 = tip_loss(t, r, a, twist[l]);
    for (int c = 0; c < 3; c++) F[c] *= F1;
    // 3D lift correction inboard of 0.75 R along the rotor axis, rotor_model.c:2548-2558
    if (t.c3d && r < 0.75 * t.r_rotor) {
        const double inflow = twist[l] * pi / 180 + a * pi / 180;
        const double Fc = -0.5 * t.c0_cl * pow(chord[l] / r, 1) * pow(cos(inflow), 2) * uref * uref * factor;
        F[0] += Fc * t.nx;
        F[1] += Fc * t.ny;
        F[2] += Fc * t.nz;
    }
    const double rn = sqrt(rx * rx + ry * ry + rz * rz);
    for (int c = 0; c < 3; c++) flagr[3 * l + c] = t.nacelle_here && rn < t.r_nacelle ? 0.0 : F[c];
}

// one pass of the force smoothing along each blade, rotor_model.c:2562-2598:
// element j in [4, ne - 4) of its blade takes sum_{i=j-4}^{j+3} F_i dfunc_s4h(i - j)
// of the previous pass (prev); the other elements keep their force
__global__ void k_alm_smooth(int nb, int ne, const double *__restrict__ prev, double *__restrict__ flagr)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nb * ne) return;
    const int b = l / ne, jj = l - b * ne;
    if (jj < 4 || jj >= ne - 4) return;
    double s[3] = {0, 0, 0};
    for (int i = l - 4; i < l + 4; i++) {
        const double w = dfunc_s4h((double)(i - l));
        for (int c = 0; c < 3; c++) s[c] += prev[3 * i + c] * w;
    }
    for (int c = 0; c < 3; c++) flagr[3 * l + c] = s[c];
}

// elements inside the nacelle radius carry no force, rotor_model.c:2599-2608
__global__ void k_alm_nacelle(TurbineDev t, const double *__restrict__ cent, double *__restrict__ flagr)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= t.nb * (t.nvb - 1)) return;
    const double rx = cent[3 * l] - t.xc, ry = cent[3 * l + 1] - t.yc, rz = cent[3 * l + 2] - t.zc;
    if (sqrt(rx * rx + ry * ry + rz * rz) < t.r_nacelle) flagr[3 * l] = flagr[3 * l + 1] = flagr[3 * l + 2] = 0.0;
}

// Calc_F_eul, rotor_model.c:3738-3830, gathered per cell over the union box
__global__ void k_alm_feul(GridMet m, int n, const double *__restrict__ cent, const int *__restrict__ box,
                           const double *__restrict__ flagr, const double *__restrict__ dA, int df, double hw, int ui0,
                           int uj0, int uk0, int uni, int unj, int unk, double *__restrict__ feul)
{
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)uni * unj * unk) return;
    const int i = ui0 + t % uni, j = uj0 + (t / uni) % unj, k = uk0 + t / ((long long)uni * unj);
    const long long q = gidx(m.g, k, j, i);
    bool geo = false;
    double ri0[3], rj0[3], rk0[3], nn[3][3], hx = 0, hy = 0, hz = 0, vol = 0;
    double acc[3] = {feul[3 * q], feul[3 * q + 1], feul[3 * q + 2]};
    for (int l = 0; l < n; l++) {
        const int *b = box + 6 * l;
        if (i < b[0] || i >= b[1] || j < b[2] || j >= b[3] || k < b[4] || k >= b[5]) continue;
        if (!geo) {
            for (int c = 0; c < 3; c++) {
                ri0[c] = (CO(m, c, k, j, i) + CO(m, c, k - 1, j, i) + CO(m, c, k, j - 1, i) + CO(m, c, k - 1, j - 1, i)) * 0.25;
                rj0[c] = (CO(m, c, k, j, i) + CO(m, c, k - 1, j, i) + CO(m, c, k, j, i - 1) + CO(m, c, k - 1, j, i - 1)) * 0.25;
                rk0[c] = (CO(m, c, k, j, i) + CO(m, c, k, j - 1, i) + CO(m, c, k, j, i - 1) + CO(m, c, k, j - 1, i - 1)) * 0.25;
            }
            cell_normals(m.csi, m.eta, m.zet, q, nn);
            cell_spacing(m.csi, m.eta, m.zet, m.aj, q, hx, hy, hz);
            vol = 1.0 / (hx * hy * hz);
            geo = true;
        }
        const double *ce = cent + 3 * l;
        const double ri[3] = {ri0[0] - ce[0], ri0[1] - ce[1], ri0[2] - ce[2]};
        const double rj[3] = {rj0[0] - ce[0], rj0[1] - ce[1], rj0[2] - ce[2]};
        const double rk[3] = {rk0[0] - ce[0], rk0[1] - ce[1], rk0[2] - ce[2]};
        double r[3][3];  // [face][direction]
        const double *rr[3] = {ri, rj, rk};
        const double h[3] = {hx, hy, hz};
        for (int f = 0; f < 3; f++)
            for (int d = 0; d < 3; d++)
                r[f][d] = fabs(rr[f][0] * nn[d][0] + rr[f][1] * nn[d][1] + rr[f][2] * nn[d][2]) / h[d];
        double dfun[3];
        for (int f = 0; f < 3; f++) {
            if (df == 0) dfun[f] = vol * dfunc_2h(r[f][0]) * dfunc_2h(r[f][1]) * dfunc_2h(r[f][2]);
            else if (df == 7) dfun[f] = vol * dfunc_exp(r[f][0], hw) * dfunc_exp(r[f][1], hw) * dfunc_exp(r[f][2], hw);
            else dfun[f] = vol * dfunc_s4h(r[f][0]) * dfunc_s4h(r[f][1]) * dfunc_s4h(r[f][2]);
        }
        const double Fx = flagr[3 * l], Fy = flagr[3 * l + 1], Fz = flagr[3 * l + 2], a = dA[l];
        const double *mt[3] = {m.csi + 3 * q, m.eta + 3 * q, m.zet + 3 * q};
        for (int f = 0; f < 3; f++)
            acc[f] += Fx * dfun[f] * a * mt[f][0] + Fy * dfun[f] * a * mt[f][1] + Fz * dfun[f] * a * mt[f][2];
    }
    for (int c = 0; c < 3; c++) feul[3 * q + c] = acc[c];
}

// zeroing after the spread, rotor_model.c:3836-3929, over the union box
// (every other cell is still zero): a cell whose 27-cell nvert sum exceeds
// 2.9 is zeroed, and so are the first and last interior layers and their
// ghosts.
__global__ void k_alm_feul_zero(GridMet m, int ui0, int uj0, int uk0, int uni, int unj, int unk, double *__restrict__ feul)
{
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)uni * unj * unk) return;
    const int i = ui0 + t % uni, j = uj0 + (t / uni) % unj, k = uk0 + t / ((long long)uni * unj);
    const Grid &g = m.g;
    if (i < 1 || i > g.mx - 2 || j < 1 || j > g.my - 2 || k < 1 || k > g.mz - 2) return;
    double nv = 0.0;
    for (int kk = k - 1; kk < k + 2; kk++)
        for (int jj = j - 1; jj < j + 2; jj++)
            for (int ii = i - 1; ii < i + 2; ii++) nv += m.nvert[gidx(g, kk, jj, ii)];
    const long long q = gidx(g, k, j, i);
    auto zero = [&](long long p) { feul[3 * p] = feul[3 * p + 1] = feul[3 * p + 2] = 0.0; };
    bool z = nv > 2.9;
    if (i == 1) { zero(gidx(g, k, j, i - 1)); z = true; }
    if (j == 1) { zero(gidx(g, k, j - 1, i)); z = true; }
    if (k == 1) { zero(gidx(g, k - 1, j, i)); z = true; }
    if (i == g.mx - 2) { zero(gidx(g, k, j, i + 1)); z = true; }
    if (j == g.my - 2) { zero(gidx(g, k, j + 1, i)); z = true; }
    if (k == g.mz - 2) { zero(gidx(g, k + 1, j, i)); z = true; }
    if (z) zero(q);
}

}  // namespace vfs

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct vfs_alm {
    vfs_ctx *ctx;
    vfs_turbine t;
    int nb, nvb, ne;
    std::vector<double> nodes, unode, cent, dA, twist, chord, refnodes, refcent, refdA;
    std::vector<int> foil_off;
    std::vector<double> foil_tab;
    std::vector<double> foil_rbeg, foil_rend;
    // device copies
    double *d_cent = nullptr, *d_refcent = nullptr, *d_dA = nullptr, *d_unode = nullptr, *d_twist = nullptr,
           *d_chord = nullptr, *d_ulagr = nullptr, *d_flagr = nullptr, *d_aoa = nullptr, *d_cl = nullptr, *d_cd = nullptr,
           *d_tab = nullptr, *d_dcent = nullptr, *d_dA_disk = nullptr, *d_ulagr_disk = nullptr;
    int *d_box = nullptr, *d_refbox = nullptr, *d_off = nullptr, *d_dbox = nullptr;
    int n_disk = 0;
    std::vector<double> disk_cent, disk_dA;
    std::vector<int> box_host;  // box of the elements spread by Calc_F_eul
    double *d_ftmp = nullptr;   // previous pass of the force smoothing
    // Calc_forces_ACL / Calc_turbineangvel state (fsi.Torque_aero, Torque_generator, ang_axis)
    double torque_aero = 0, torque_generator = 0, ang_axis = 0;
};

namespace vfs {

static vfs::GridMet gridmet(vfs_ctx *c)
{
    GridMet m;
    m.g = c->g;
    m.coor = c->f[VFS_COOR];
    m.csi = c->f[VFS_CSI];
    m.eta = c->f[VFS_ETA];
    m.zet = c->f[VFS_ZET];
    m.aj = c->f[VFS_AJ];
    m.nvert = c->f[VFS_NVERT];
    return m;
}

// ArbitraryRotate, rotor_model.c:5174-5192
static void arbitrary_rotate(const double p[3], double theta, const double r_in[3], double q[3])
{
#pragma clang fp contract(off)
    double r[3] = {r_in[0], r_in[1], r_in[2]};
    const double rr = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) + 1.0e-11;
    r[0] = r[0] / rr; r[1] = r[1] / rr; r[2] = r[2] / rr;
    const double c = std::cos(theta), s = std::sin(theta);
    q[0] = q[1] = q[2] = 0.0;
    q[0] += (c + (1 - c) * r[0] * r[0]) * p[0];
    q[0] += ((1 - c) * r[0] * r[1] - r[2] * s) * p[1];
    q[0] += ((1 - c) * r[0] * r[2] + r[1] * s) * p[2];
    q[1] += ((1 - c) * r[0] * r[1] + r[2] * s) * p[0];
    q[1] += (c + (1 - c) * r[1] * r[1]) * p[1];
    q[1] += ((1 - c) * r[1] * r[2] - r[0] * s) * p[2];
    q[2] += ((1 - c) * r[0] * r[2] - r[1] * s) * p[0];
    q[2] += ((1 - c) * r[1] * r[2] + r[0] * s) * p[1];
    q[2] += (c + (1 - c) * r[2] * r[2]) * p[2];
}

// element centres and lengths of line elements (rotor_model.c:4838-4864)
static void line_elements(int nb, int nvb, const std::vector<double> &x, std::vector<double> &cent, std::vector<double> &dA)
{
#pragma clang fp contract(off)
    const int ne = nvb - 1;
    for (int b = 0; b < nb; b++)
        for (int j = 0; j < ne; j++) {
            const int i = b * ne + j, n1 = b * nvb + j, n2 = n1 + 1;
            const double dx = x[3 * n2] - x[3 * n1], dy = x[3 * n2 + 1] - x[3 * n1 + 1], dz = x[3 * n2 + 2] - x[3 * n1 + 2];
            dA[i] = std::sqrt(dx * dx + dy * dy + dz * dz);
            for (int c = 0; c < 3; c++) cent[3 * i + c] = (x[3 * n1 + c] + x[3 * n2 + c]) / 2.;
        }
}

// twist and chord by radius (airfoil_ACL, rotor_model.c:2118-2134)
static void twist_chord(vfs_alm *a)
{
#pragma clang fp contract(off)
    const vfs_turbine &t = a->t;
    for (int i = 0; i < a->nb * a->ne; i++) {
        const double r = std::sqrt(std::pow(a->cent[3 * i] - t.xc, 2.0) + std::pow(a->cent[3 * i + 1] - t.yc, 2.0) +
                                   std::pow(a->cent[3 * i + 2] - t.zc, 2.0));
        for (size_t f = 0; f < a->foil_rbeg.size(); f++) {
            if (!(r >= a->foil_rbeg[f] && r <= a->foil_rend[f])) continue;
            const int *o = &a->foil_off[6 * f];
            const double *rac = &a->foil_tab[o[1]], *tw = rac + o[0], *ch = rac + 2 * o[0];
            for (int j = 0; j < o[0] - 1; j++)
                if (r >= rac[j] && r <= rac[j + 1]) {
                    const double fac1 = (rac[j + 1] - r) / (rac[j + 1] - rac[j]);
                    const double fac2 = (-rac[j] + r) / (rac[j + 1] - rac[j]);
                    a->twist[i] = fac1 * tw[j] + fac2 * tw[j + 1];
                    a->chord[i] = fac1 * ch[j] + fac2 * ch[j + 1];
                }
        }
    }
}

static int upload(vfs_alm *a, double *dst, const std::vector<double> &src)
{
    VFS_HIP(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(double), hipMemcpyHostToDevice, a->ctx->stream));
    return VFS_OK;
}

// reference line of rotor_model 6 (refAL_Rot, rotor_model.c:5034-5088)
static void ref_line(vfs_alm *a)
{
#pragma clang fp contract(off)
    const vfs_turbine &t = a->t;
    const double na[3] = {t.nx, t.ny, t.nz};
    double isign = (1.e-19 + t.angvel) / (std::fabs(t.angvel) + 1.e-19);
    isign = isign > 0.0 ? 1.0 : -1.0;
    const double theta = isign * t.refangle * M_PI / 180.0;
    for (int i = 0; i < a->nb * a->nvb; i++) {
        const double p[3] = {a->nodes[3 * i] - t.xc, a->nodes[3 * i + 1] - t.yc, a->nodes[3 * i + 2] - t.zc};
        double q[3];
        arbitrary_rotate(p, theta, na, q);
        a->refnodes[3 * i] = q[0] + t.xc;
        a->refnodes[3 * i + 1] = q[1] + t.yc;
        a->refnodes[3 * i + 2] = q[2] + t.zc;
    }
    line_elements(a->nb, a->nvb, a->refnodes, a->refcent, a->refdA);
}

}  // namespace vfs

using namespace vfs;

extern "C" {

int vfs_alm_create(vfs_ctx *c, const vfs_turbine *t, int num_blade, int nv_blade, const double *nodes, int nfoil,
                   const vfs_foil *foils, vfs_alm **out)
{
    if (!c || !t || !nodes || !out || num_blade < 1 || nv_blade < 2 || nfoil < 0 || (nfoil > 0 && !foils))
        return set_error(VFS_EINVAL, "vfs_alm_create: bad arguments");
    if (t->rotor_model != 3 && t->rotor_model != 6)
        return set_error(VFS_EINVAL, "vfs_alm_create: rotor_model must be 3 or 6 (actuator line)");
    vfs_alm *a = new vfs_alm();
    a->ctx = c;
    a->t = *t;
    const double rr = std::sqrt(t->nx * t->nx + t->ny * t->ny + t->nz * t->nz);  // main.c:2051-2054
    if (!(rr > 0)) { delete a; return set_error(VFS_EINVAL, "vfs_alm_create: zero rotor axis"); }
    a->t.nx = t->nx / rr; a->t.ny = t->ny / rr; a->t.nz = t->nz / rr;
    a->nb = num_blade;
    a->nvb = nv_blade;
    a->ne = nv_blade - 1;
    const int nv = num_blade * nv_blade, ne = num_blade * a->ne;
    a->nodes.assign(nodes, nodes + 3 * nv);
    a->unode.assign(3 * nv, 0.0);
    a->refnodes.assign(3 * nv, 0.0);
    a->cent.assign(3 * ne, 0.0); a->dA.assign(ne, 0.0);
    a->refcent.assign(3 * ne, 0.0); a->refdA.assign(ne, 0.0);
    a->twist.assign(ne, 0.0); a->chord.assign(ne, 0.0);
    for (int f = 0; f < nfoil; f++) {
        const vfs_foil &F = foils[f];
        if (F.n_ac < 1 || F.n_cl < 1 || F.n_cd < 1 || !F.r_ac || !F.twist || !F.chord || !F.ang_cl || !F.cl || !F.ang_cd || !F.cd) {
            delete a;
            return set_error(VFS_EINVAL, "vfs_alm_create: incomplete foil table");
        }
        const int o_ac = (int)a->foil_tab.size();
        a->foil_tab.insert(a->foil_tab.end(), F.r_ac, F.r_ac + F.n_ac);
        a->foil_tab.insert(a->foil_tab.end(), F.twist, F.twist + F.n_ac);
        a->foil_tab.insert(a->foil_tab.end(), F.chord, F.chord + F.n_ac);
        const int o_cl = (int)a->foil_tab.size();
        a->foil_tab.insert(a->foil_tab.end(), F.ang_cl, F.ang_cl + F.n_cl);
        a->foil_tab.insert(a->foil_tab.end(), F.cl, F.cl + F.n_cl);
        const int o_cd = (int)a->foil_tab.size();
        a->foil_tab.insert(a->foil_tab.end(), F.ang_cd, F.ang_cd + F.n_cd);
        a->foil_tab.insert(a->foil_tab.end(), F.cd, F.cd + F.n_cd);
        const int o[6] = {F.n_ac, o_ac, F.n_cl, o_cl, F.n_cd, o_cd};
        a->foil_off.insert(a->foil_off.end(), o, o + 6);
        a->foil_rbeg.push_back(F.r_ac[0]);
        a->foil_rend.push_back(F.r_ac[F.n_ac - 1]);
    }
    line_elements(a->nb, a->nvb, a->nodes, a->cent, a->dA);
    twist_chord(a);
    auto dal = [&](void **p, size_t bytes) -> int {
        hipError_t e = hipMalloc(p, bytes ? bytes : 8);
        if (e != hipSuccess) return check_hip(e, "hipMalloc (alm)");
        e = hipMemset(*p, 0, bytes ? bytes : 8);
        return e == hipSuccess ? VFS_OK : check_hip(e, "hipMemset (alm)");
    };
    int rc = 0;
    rc = rc ? rc : dal((void **)&a->d_cent, 3 * ne * 8);
    rc = rc ? rc : dal((void **)&a->d_refcent, 3 * ne * 8);
    rc = rc ? rc : dal((void **)&a->d_dA, ne * 8);
    rc = rc ? rc : dal((void **)&a->d_unode, 3 * nv * 8);
    rc = rc ? rc : dal((void **)&a->d_twist, ne * 8);
    rc = rc ? rc : dal((void **)&a->d_chord, ne * 8);
    rc = rc ? rc : dal((void **)&a->d_ulagr, 3 * ne * 8);
    rc = rc ? rc : dal((void **)&a->d_flagr, 3 * ne * 8);
    rc = rc ? rc : dal((void **)&a->d_ftmp, 3 * ne * 8);
    rc = rc ? rc : dal((void **)&a->d_aoa, ne * 8);
    rc = rc ? rc : dal((void **)&a->d_cl, ne * 8);
    rc = rc ? rc : dal((void **)&a->d_cd, ne * 8);
    rc = rc ? rc : dal((void **)&a->d_box, 6 * ne * sizeof(int));
    rc = rc ? rc : dal((void **)&a->d_refbox, 6 * ne * sizeof(int));
    rc = rc ? rc : dal((void **)&a->d_tab, a->foil_tab.size() * 8);
    rc = rc ? rc : dal((void **)&a->d_off, a->foil_off.size() * sizeof(int));
    if (!rc && !a->foil_tab.empty())
        rc = hipMemcpy(a->d_tab, a->foil_tab.data(), a->foil_tab.size() * 8, hipMemcpyHostToDevice) == hipSuccess ? 0 : check_hip(hipErrorUnknown, "foil upload");
    if (!rc && !a->foil_off.empty())
        rc = hipMemcpy(a->d_off, a->foil_off.data(), a->foil_off.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess ? 0 : check_hip(hipErrorUnknown, "foil upload");
    if (!rc) rc = upload(a, a->d_cent, a->cent);
    if (!rc) rc = upload(a, a->d_dA, a->dA);
    if (!rc) rc = upload(a, a->d_twist, a->twist);
    if (!rc) rc = upload(a, a->d_chord, a->chord);
    if (rc) { vfs_alm_destroy(a); return rc; }
    *out = a;
    return VFS_OK;
}

int vfs_alm_destroy(vfs_alm *a)
{
    if (!a) return VFS_OK;
    hipStreamSynchronize(a->ctx->stream);
    void *ps[] = {a->d_cent, a->d_refcent, a->d_dA, a->d_unode, a->d_twist, a->d_chord, a->d_ulagr, a->d_flagr, a->d_aoa,
                  a->d_cl, a->d_cd, a->d_tab, a->d_box, a->d_refbox, a->d_off, a->d_dcent, a->d_dA_disk, a->d_ulagr_disk,
                  a->d_dbox, a->d_ftmp};
    for (void *p : ps)
        if (p) hipFree(p);
    delete a;
    return VFS_OK;
}

int vfs_alm_rotate(vfs_alm *a, double dt)
{
#pragma clang fp contract(off)
    if (!a) return set_error(VFS_EINVAL, "vfs_alm_rotate: null handle");
    const vfs_turbine &t = a->t;
    const double na[3] = {t.nx, t.ny, t.nz};
    for (int i = 0; i < a->nb * a->nvb; i++) {  // rotor_Rot, rotor_model.c:4775-4806
        const double p[3] = {a->nodes[3 * i] - t.xc, a->nodes[3 * i + 1] - t.yc, a->nodes[3 * i + 2] - t.zc};
        double q[3];
        arbitrary_rotate(p, t.angvel * dt, na, q);
        a->nodes[3 * i] = q[0] + t.xc;
        a->nodes[3 * i + 1] = q[1] + t.yc;
        a->nodes[3 * i + 2] = q[2] + t.zc;
        const double rx = a->nodes[3 * i] - t.xc, ry = a->nodes[3 * i + 1] - t.yc, rz = a->nodes[3 * i + 2] - t.zc;
        const double rr = std::sqrt(rx * rx + ry * ry + rz * rz) + 1.e-19;
        const double nr[3] = {rx / rr, ry / rr, rz / rr};
        const double nt[3] = {na[1] * nr[2] - na[2] * nr[1], na[2] * nr[0] - na[0] * nr[2], na[0] * nr[1] - na[1] * nr[0]};
        const double ut = t.angvel * rr;
        for (int c = 0; c < 3; c++) a->unode[3 * i + c] = ut * nt[c];
    }
    line_elements(a->nb, a->nvb, a->nodes, a->cent, a->dA);
    int rc = upload(a, a->d_cent, a->cent);
    if (!rc) rc = upload(a, a->d_dA, a->dA);
    if (!rc) rc = upload(a, a->d_unode, a->unode);
    if (!rc && t.rotor_model == 6) {
        ref_line(a);
        rc = upload(a, a->d_refcent, a->refcent);
    }
    return rc;
}

int vfs_alm_preprocess(vfs_alm *a)
{
    if (!a) return set_error(VFS_EINVAL, "vfs_alm_preprocess: null handle");
    vfs_ctx *c = a->ctx;
    const int ne = a->nb * a->ne;
    GridMet m = gridmet(c);
    hipLaunchKernelGGL(k_alm_nearest, dim3(blocks_for(ne, 64)), dim3(64), 0, c->stream, m, ne, a->d_cent, a->d_box);
    if (a->t.rotor_model == 6)
        hipLaunchKernelGGL(k_alm_nearest, dim3(blocks_for(ne, 64)), dim3(64), 0, c->stream, m, ne, a->d_refcent, a->d_refbox);
    VFS_LAUNCH_CHECK("alm nearest");
    a->box_host.resize(6 * ne);
    VFS_HIP(hipMemcpyAsync(a->box_host.data(), a->d_box, 6 * ne * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    VFS_HIP(hipStreamSynchronize(c->stream));
    return VFS_OK;
}

int vfs_alm_u_lagr(vfs_alm *a)
{
    if (!a) return set_error(VFS_EINVAL, "vfs_alm_u_lagr: null handle");
    vfs_ctx *c = a->ctx;
    const int ne = a->nb * a->ne;
    const bool ref = a->t.rotor_model == 6;  // model 6 samples on the reference line (solvers.c:543-555)
    hipLaunchKernelGGL(k_alm_ulagr, dim3(ne), dim3(ALM_TPB), 0, c->stream, gridmet(c), (const double *)c->f[VFS_UCAT], ne,
                       ref ? a->d_refcent : a->d_cent, ref ? a->d_refbox : a->d_box, a->d_ulagr);
    VFS_LAUNCH_CHECK("alm u_lagr");
    return VFS_OK;
}

int vfs_alm_f_lagr(vfs_alm *a)
{
    if (!a) return set_error(VFS_EINVAL, "vfs_alm_f_lagr: null handle");
    vfs_ctx *c = a->ctx;
    const int ne = a->nb * a->ne;
    const vfs_turbine &T = a->t;
    const int smooth = T.smoothforce;
    TurbineDev td{T.nx, T.ny, T.nz, T.xc, T.yc, T.zc, T.pitch, T.r_nacelle, a->nb, a->nvb, T.shen, T.prandtl, T.shen1,
                  T.correction3d_cl, T.num_blade, smooth ? 0 : 1, T.r_rotor, T.tsr, T.a_shen, T.b_shen, T.c_shen,
                  T.correction_shen1, T.c0_cl};
    Polars P{(int)a->foil_rbeg.size(), a->d_off, a->d_tab};
    hipLaunchKernelGGL(k_alm_flagr, dim3(blocks_for(ne, 64)), dim3(64), 0, c->stream, td, P, a->d_cent, a->d_ulagr, a->d_unode,
                       a->d_twist, a->d_chord, a->d_aoa, a->d_cl, a->d_cd, a->d_flagr);
    if (smooth) {
        // smoothforce_AL + 1 passes (the reference's do-while), then the nacelle zeroing
        for (int pass = 0; pass <= smooth; pass++) {
            VFS_HIP(hipMemcpyAsync(a->d_ftmp, a->d_flagr, 3 * (size_t)ne * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            hipLaunchKernelGGL(k_alm_smooth, dim3(blocks_for(ne, 64)), dim3(64), 0, c->stream, a->nb, a->ne, a->d_ftmp,
                               a->d_flagr);
        }
        hipLaunchKernelGGL(k_alm_nacelle, dim3(blocks_for(ne, 64)), dim3(64), 0, c->stream, td, a->d_cent, a->d_flagr);
    }
    VFS_LAUNCH_CHECK("alm f_lagr");
    return VFS_OK;
}

int vfs_alm_loads(vfs_alm *a, double *out)
{
    // Calc_forces_ACL, rotor_model.c:4148-4180, summed on the host in element order
    if (!a || !out) return set_error(VFS_EINVAL, "vfs_alm_loads: null argument");
    vfs_ctx *c = a->ctx;
    const int ne = a->nb * a->ne;
    std::vector<double> F(3 * (size_t)ne), U(3 * (size_t)ne);
    VFS_HIP(hipMemcpyAsync(F.data(), a->d_flagr, F.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    VFS_HIP(hipMemcpyAsync(U.data(), a->d_ulagr, U.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    VFS_HIP(hipStreamSynchronize(c->stream));
    double A_Sum = 0, F_Sum = 0, P_Sum = 0, U_Sum = 0, Fx = 0, Fy = 0, Fz = 0, M_Sum = 0;
    const double nx = a->t.nx, ny = a->t.ny, nz = a->t.nz;
    for (int l = 0; l < ne; l++) {
        const double *f = &F[3 * l], *u = &U[3 * l], dA = a->dA[l];
        A_Sum += dA;
        const double F_axis = f[0] * nx + f[1] * ny + f[2] * nz;
        const double U_axis = u[0] * nx + u[1] * ny + u[2] * nz;
        F_Sum += F_axis * dA;
        P_Sum += F_axis * dA * U_axis;
        U_Sum += U_axis * dA;
        Fx += f[0] * dA;
        Fy += f[1] * dA;
        Fz += f[2] * dA;
        const double rx = a->cent[3 * l] - a->t.xc, ry = a->cent[3 * l + 1] - a->t.yc, rz = a->cent[3 * l + 2] - a->t.zc;
        const double M_x = ry * f[2] * dA - rz * f[1] * dA;
        const double M_y = rz * f[0] * dA - rx * f[2] * dA;
        const double M_z = rx * f[1] * dA - ry * f[0] * dA;
        M_Sum += M_x * nx + M_y * ny + M_z * nz;
    }
    const double v[8] = {A_Sum, F_Sum, P_Sum, U_Sum / A_Sum, Fx, Fy, Fz, M_Sum};
    std::memcpy(out, v, sizeof(v));
    a->torque_aero = -M_Sum;
    return VFS_OK;
}

int vfs_alm_turbine_angvel(vfs_alm *a, const vfs_turbine_control *tc, double dt, double uref, int ti, int tistart,
                           double *angvel, double *tsr)
{
    // Calc_turbineangvel, rotor_model.c:5519-5601 (one turbine; no restart file)
#pragma clang fp contract(off)
    if (!a || !tc) return set_error(VFS_EINVAL, "vfs_alm_turbine_angvel: null argument");
    if (tc->mode < 0 || tc->mode > 2) return set_error(VFS_EINVAL, "vfs_alm_turbine_angvel: mode must be 0, 1 or 2");
    if (!(tc->reflength_wt > 0) || !(uref != 0)) return set_error(VFS_EINVAL, "vfs_alm_turbine_angvel: reflength_wt and uref");
    const double R0 = tc->r_rotor / tc->reflength_wt;
    const double rho_cfd = 1.0;
    const double T_wt = tc->r_rotor / tc->refvel_wt;
    const double T_cfd = R0 / tc->refvel_cfd;
    const double ratio_rho = tc->rho_air / rho_cfd, ratio_L = tc->reflength_wt, ratio_T = T_wt / T_cfd;
    const double ratio_V = tc->refvel_wt / tc->refvel_cfd;
    const double R = tc->r_rotor / tc->reflength_wt;
    vfs_turbine &T = a->t;
    if (ti == tistart) {
        a->ang_axis = 0.0;
        if (tc->mode == 1) T.angvel = tc->tsr * uref / R, T.tsr = tc->tsr;
        else if (tc->mode == 2) T.angvel = tc->TSR_max * uref / R;
        else T.angvel = tc->angvel_fixed, T.tsr = T.angvel * R / uref;
    } else {
        if (tc->mode == 1) {
            T.tsr = tc->tsr;
            T.angvel = T.tsr * uref / R;
        } else if (tc->mode == 0) {
            T.angvel = tc->angvel_fixed;
            T.tsr = T.angvel * R / uref;
        } else {
            const double J = tc->J_rotation;
            const double K = 0.5 * M_PI * pow(tc->r_rotor, 2 + tc->X_control) * tc->CP_max / pow(tc->TSR_max, tc->X_control);
            const double angvel_wt = T.angvel / ratio_T;
            const double UU = uref * ratio_V;
            if (!tc->fixturbinegeneratortorque) a->torque_generator = K * pow(angvel_wt, tc->X_control - 1) * pow(UU, 3 - tc->X_control);
            else a->torque_generator = tc->torque_generator;
            const double Torque_c = a->torque_generator;
            const double Torque_a = a->torque_aero * ratio_rho * pow(ratio_V, 2) * pow(ratio_L, 3);
            const double Torque = Torque_a - Torque_c;
            const double dt_wt = dt * ratio_T;
            const double angvel_0 = T.angvel / ratio_T;
            const double angvel_1 = angvel_0 + dt_wt * Torque / J;
            T.angvel = angvel_1 * ratio_T;
            T.tsr = T.angvel * R / uref;
        }
    }
    a->ang_axis += T.angvel * dt;
    if (angvel) *angvel = T.angvel;
    if (tsr) *tsr = T.tsr;
    return VFS_OK;
}

int vfs_alm_f_eul(vfs_alm *a)
{
    if (!a) return set_error(VFS_EINVAL, "vfs_alm_f_eul: null handle");
    vfs_ctx *c = a->ctx;
    const int ne = a->nb * a->ne;
    if ((int)a->box_host.size() != 6 * ne) return set_error(VFS_EINVAL, "vfs_alm_f_eul: call vfs_alm_preprocess first");
    // VecSet(lF_eul, 0), solvers.c:484
    VFS_HIP(hipMemsetAsync(c->f[VFS_FEUL], 0, (size_t)3 * c->g.n * sizeof(double), c->stream));
    int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-1, -1, -1};
    for (int l = 0; l < ne; l++)
        for (int d = 0; d < 3; d++) {
            const int b0 = a->box_host[6 * l + 2 * d], b1 = a->box_host[6 * l + 2 * d + 1];
            if (b1 <= b0) continue;
            lo[d] = std::min(lo[d], b0);
            hi[d] = std::max(hi[d], b1);
        }
    if (hi[0] < 0 || hi[1] < 0 || hi[2] < 0) return VFS_OK;  // nothing to spread
    const int uni = hi[0] - lo[0], unj = hi[1] - lo[1], unk = hi[2] - lo[2];
    const long long cells = (long long)uni * unj * unk;
    GridMet m = gridmet(c);
    hipLaunchKernelGGL(k_alm_feul, dim3(blocks_for(cells, ALM_TPB)), dim3(ALM_TPB), 0, c->stream, m, ne, a->d_cent, a->d_box,
                       a->d_flagr, a->d_dA, a->t.deltafunc, a->t.halfwidth, lo[0], lo[1], lo[2], uni, unj, unk,
                       c->f[VFS_FEUL]);
    hipLaunchKernelGGL(k_alm_feul_zero, dim3(blocks_for(cells, ALM_TPB)), dim3(ALM_TPB), 0, c->stream, m, lo[0], lo[1], lo[2],
                       uni, unj, unk, c->f[VFS_FEUL]);
    VFS_LAUNCH_CHECK("alm f_eul");
    return VFS_OK;
}

int vfs_alm_step(vfs_alm *a, double dt)
{
    // solvers.c:535-563 (rotor_model 6) / 497-524 with rotor_Rot (rotor_model 3), fixed angular velocity
    int rc = vfs_alm_rotate(a, dt);
    if (!rc) rc = vfs_alm_preprocess(a);
    if (!rc) rc = vfs_alm_u_lagr(a);
    if (!rc) rc = vfs_alm_f_lagr(a);
    if (!rc) rc = vfs_alm_f_eul(a);
    return rc;
}

int vfs_alm_get(vfs_alm *a, vfs_alm_array what, double *host)
{
    if (!a || !host) return set_error(VFS_EINVAL, "vfs_alm_get: bad arguments");
    vfs_ctx *c = a->ctx;
    const int ne = a->nb * a->ne, nv = a->nb * a->nvb;
    auto dl = [&](const double *d, size_t n) -> int {
        VFS_HIP(hipMemcpyAsync(host, d, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        VFS_HIP(hipStreamSynchronize(c->stream));
        return VFS_OK;
    };
    switch (what) {
    case VFS_ALM_NODES: std::memcpy(host, a->nodes.data(), 3 * nv * sizeof(double)); return VFS_OK;
    case VFS_ALM_NODE_VEL: std::memcpy(host, a->unode.data(), 3 * nv * sizeof(double)); return VFS_OK;
    case VFS_ALM_CENT: std::memcpy(host, a->cent.data(), 3 * ne * sizeof(double)); return VFS_OK;
    case VFS_ALM_REF_CENT: std::memcpy(host, a->refcent.data(), 3 * ne * sizeof(double)); return VFS_OK;
    case VFS_ALM_DA: std::memcpy(host, a->dA.data(), ne * sizeof(double)); return VFS_OK;
    case VFS_ALM_TWIST: std::memcpy(host, a->twist.data(), ne * sizeof(double)); return VFS_OK;
    case VFS_ALM_CHORD: std::memcpy(host, a->chord.data(), ne * sizeof(double)); return VFS_OK;
    case VFS_ALM_U_LAGR: return dl(a->d_ulagr, 3 * ne);
    case VFS_ALM_F_LAGR: return dl(a->d_flagr, 3 * ne);
    case VFS_ALM_AOA: return dl(a->d_aoa, ne);
    case VFS_ALM_CL: return dl(a->d_cl, ne);
    case VFS_ALM_CD: return dl(a->d_cd, ne);
    case VFS_ALM_BOX:
    case VFS_ALM_REF_BOX: {
        std::vector<int> b(6 * ne);
        VFS_HIP(hipMemcpyAsync(b.data(), what == VFS_ALM_BOX ? a->d_box : a->d_refbox, 6 * ne * sizeof(int),
                               hipMemcpyDeviceToHost, c->stream));
        VFS_HIP(hipStreamSynchronize(c->stream));
        for (int q = 0; q < 6 * ne; q++) host[q] = b[q];
        return VFS_OK;
    }
    }
    return set_error(VFS_EINVAL, "vfs_alm_get: unknown array");
}

int vfs_alm_counts(const vfs_alm *a, int *n_elements, int *n_nodes)
{
    if (!a) return set_error(VFS_EINVAL, "vfs_alm_counts: null handle");
    if (n_elements) *n_elements = a->nb * a->ne;
    if (n_nodes) *n_nodes = a->nb * a->nvb;
    return VFS_OK;
}

int vfs_alm_set_disk(vfs_alm *a, int n, const double *cent, const double *dA)
{
    if (!a || n < 1 || !cent || !dA) return set_error(VFS_EINVAL, "vfs_alm_set_disk: bad arguments");
    vfs_ctx *c = a->ctx;
    void *ps[] = {a->d_dcent, a->d_dA_disk, a->d_ulagr_disk, a->d_dbox};
    hipStreamSynchronize(c->stream);
    for (void *p : ps)
        if (p) hipFree(p);
    a->n_disk = n;
    a->disk_cent.assign(cent, cent + 3 * n);
    a->disk_dA.assign(dA, dA + n);
    VFS_HIP(hipMalloc((void **)&a->d_dcent, 3 * n * 8));
    VFS_HIP(hipMalloc((void **)&a->d_dA_disk, n * 8));
    VFS_HIP(hipMalloc((void **)&a->d_ulagr_disk, 3 * n * 8));
    VFS_HIP(hipMalloc((void **)&a->d_dbox, 6 * n * sizeof(int)));
    VFS_HIP(hipMemcpy(a->d_dcent, cent, 3 * n * 8, hipMemcpyHostToDevice));
    VFS_HIP(hipMemcpy(a->d_dA_disk, dA, n * 8, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_alm_nearest, dim3(blocks_for(n, 64)), dim3(64), 0, c->stream, gridmet(c), n, a->d_dcent, a->d_dbox);
    VFS_LAUNCH_CHECK("alm disk nearest");
    VFS_HIP(hipStreamSynchronize(c->stream));
    return VFS_OK;
}

int vfs_alm_uref(vfs_alm *a, double *uref)
{
#pragma clang fp contract(off)
    if (!a || !uref || a->n_disk < 1) return set_error(VFS_EINVAL, "vfs_alm_uref: no reference disk (vfs_alm_set_disk)");
    vfs_ctx *c = a->ctx;
    const int n = a->n_disk;
    hipLaunchKernelGGL(k_alm_ulagr, dim3(n), dim3(ALM_TPB), 0, c->stream, gridmet(c), (const double *)c->f[VFS_UCAT], n,
                       a->d_dcent, a->d_dbox, a->d_ulagr_disk);
    VFS_LAUNCH_CHECK("alm disk u_lagr");
    std::vector<double> u(3 * n);
    VFS_HIP(hipMemcpyAsync(u.data(), a->d_ulagr_disk, 3 * n * 8, hipMemcpyDeviceToHost, c->stream));
    VFS_HIP(hipStreamSynchronize(c->stream));
    double U = 0.0, A = 0.0;  // Uref_ACL, rotor_model.c:4385-4393
    for (int l = 0; l < n; l++) {
        const double ua = u[3 * l] * a->t.nx + u[3 * l + 1] * a->t.ny + u[3 * l + 2] * a->t.nz;
        U += ua * a->disk_dA[l];
        A += a->disk_dA[l];
    }
    *uref = U / A;
    return VFS_OK;
}

}  // extern "C"
