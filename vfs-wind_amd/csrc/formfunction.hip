// formfunction.hip -- Formfunction_2 (momentum.c:454-2013) for gfx950.
//
// Two passes over the DMDA block:
//   face pass: one thread per point computes the convective and viscous
//     fluxes of the i, j and k faces it owns (momentum.c:668-1451) and
//     writes them as 18 structure-of-arrays planes (coalesced stores);
//   cell pass: one thread per point forms the flux divergence Fp of its cell
//     and of the three upper neighbours (momentum.c:1548-1678), projects
//     them on the contravariant directions (momentum.c:1715-1841), applies
//     the boundary zeroing (momentum.c:1833-1938) and either accumulates into
//     Rhs or, inside RungeKutta, writes Ucont = a*(Rhs - dP) + pUcont
//     directly (rhs.c:1647-1658), so Rhs never round-trips through HBM.
// The arithmetic follows the reference's operation order, with FP
// contraction disabled for this file, so results match the FP64 CPU path to
// a few ulps.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "vfs_internal.h"
#include "vfs_reduce.h"

#pragma clang fp contract(off)

namespace vfs {

constexpr int FF_TPB = 256;

struct Metrics9 {
    const double *csi, *eta, *zet, *aj;  // one face family (I, J or K)
};

struct FaceArgs {
    Grid g;
    const double *ucont, *ucat, *nut, *nvert;
    Metrics9 fi, fj, fk;
    Metrics9 cell;  // cell metrics csi, eta, zet, aj
    double nu;
    int les, second_order, laplacian;
    double *out;  // 18 planes of n doubles: div1,div2,div3 (xyz each), visc1..3
};

// ucat, nvert and nu_t of the points around the block's current plane come
// from an LDS tile of 3 k-planes around the block's FT_I x FT_J points.  With
// the second-order convection (-second_order 1) every face of the block
// reads only the points i-1..i+1, j-1..j+1, k-1..k+1 of its own point, so the
// tile spans i0-1..i0+FT_I and j0-1..j0+FT_J and the accessor reads LDS only.
// The fourth-order QUICK path also reads i+2 / j+2 (one more tile column and
// row), the k+2 plane and, at a periodic boundary, the far ghosts; those
// reads outside the tile go to HBM.  The tile holds exactly the values the
// DMDA index with periodic wrap would return.  nvert (blanking codes, small
// integers) is kept as float in LDS, which represents them exactly.
constexpr int FT_I = 32, FT_J = 16, FT_K = 16;

template <bool SO>
struct TileDims {
    static constexpr int HI = FT_I + (SO ? 2 : 3), HJ = FT_J + (SO ? 2 : 3), HP = HI * HJ;
};

// NB: no point of the grid is blanked (every nvert is 0), so NV is the
// constant 0 and every blanking test folds away at compile time; the tests
// then take the branches they take for nvert = 0, so the result is the same.
template <bool SO, bool NB = false>
struct TileAcc {
    using D = TileDims<SO>;
    const FaceArgs *a;
    const double *ux, *uy, *uz, *nt;
    const float *nv;
    int i0, j0, kc, slot0;  // slot0 holds plane kc-1
    __device__ inline int lidx(int k, int j, int i, bool &in) const
    {
        const int di = i - i0 + 1, dj = j - j0 + 1, dk = k - kc + 1;
        in = SO || ((unsigned)di < (unsigned)D::HI && (unsigned)dj < (unsigned)D::HJ && (unsigned)dk < 3u);
        int s = slot0 + dk;
        if (s >= 3) s -= 3;
        return (s * D::HJ + dj) * D::HI + di;
    }
    __device__ inline double U(int c, int k, int j, int i) const
    {
        bool in;
        const int l = lidx(k, j, i, in);
        if (in) return c == 0 ? ux[l] : (c == 1 ? uy[l] : uz[l]);
        return a->ucat[3 * gidx(a->g, k, j, i) + c];
    }
    __device__ inline double NV(int k, int j, int i) const
    {
        if constexpr (NB) return 0.0;
        bool in;
        const int l = lidx(k, j, i, in);
        return in ? (double)nv[l] : a->nvert[gidx(a->g, k, j, i)];
    }
    __device__ inline double NT(int k, int j, int i) const
    {
        bool in;
        const int l = lidx(k, j, i, in);
        return in ? nt[l] : a->nut[gidx(a->g, k, j, i)];
    }
};

// viscous face flux, momentum.c:682-696 and 863-902 (and the j/k twins)
__device__ inline void visc_face(const double cs[3], const double et[3], const double ze[3], const double nrm[3],
                                 double ajc, const double d[3][3], double nu, double nu_t, int les, int laplacian,
                                 double out[3])
{
    double g11 = cs[0] * nrm[0] + cs[1] * nrm[1] + cs[2] * nrm[2];
    double g21 = et[0] * nrm[0] + et[1] * nrm[1] + et[2] * nrm[2];
    double g31 = ze[0] * nrm[0] + ze[1] * nrm[1] + ze[2] * nrm[2];
    double r[3][3];
    for (int c = 0; c < 3; c++)
        for (int q = 0; q < 3; q++) r[c][q] = d[c][0] * cs[q] + d[c][1] * et[q] + d[c][2] * ze[q];
    for (int c = 0; c < 3; c++) out[c] = 0;
    if (les)
        for (int c = 0; c < 3; c++)
            out[c] = (g11 * d[c][0] + g21 * d[c][1] + g31 * d[c][2] + r[0][c] * nrm[0] + r[1][c] * nrm[1] +
                      r[2][c] * nrm[2]) * ajc * (nu_t);
    if (laplacian)
        for (int c = 0; c < 3; c++)
            for (int q = 0; q < 3; q++) r[c][q] = 0.;
    for (int c = 0; c < 3; c++)
        out[c] += (g11 * d[c][0] + g21 * d[c][1] + g31 * d[c][2] + r[0][c] * nrm[0] + r[1][c] * nrm[1] +
                   r[2][c] * nrm[2]) * ajc * (nu);
}

// the 10 metric values (csi, eta, zet, aj) of point q of one family
__device__ inline void load_metrics(const Metrics9 &m, long long q, double v[10])
{
#pragma unroll
    for (int c = 0; c < 3; c++) {
        v[c] = m.csi[3 * q + c];
        v[3 + c] = m.eta[3 * q + c];
        v[6 + c] = m.zet[3 * q + c];
    }
    v[9] = m.aj[q];
}

// convective and viscous flux of the i face of point (i,j,k)
// (momentum.c:668-928): i in [0,mx-2], j,k interior; zero elsewhere
template <bool SO, bool NB = false>
__device__ inline void flux_i(const FaceArgs &a, const TileAcc<SO, NB> &T, int i, int j, int k, const double *Mf, double ucf,
                              double dv[3], double vs[3])
{
    const Grid &g = a.g;
    const int mx = g.mx, my = g.my, mz = g.mz;
    const double SOL = 0.1;  // k-omega.c:19
    (void)SOL; (void)my; (void)mz;
    for (int c = 0; c < 3; c++) dv[c] = vs[c] = 0;
    if (i < mx - 1 && j >= 1 && j < my - 1 && k >= 1 && k < mz - 1) {
        const double cs[3] = {Mf[0], Mf[1], Mf[2]};
        const double et[3] = {Mf[3], Mf[4], Mf[5]};
        const double ze[3] = {Mf[6], Mf[7], Mf[8]};
        double d[3][3];
        // Compute_du_i, k-omega.c:56-116
        for (int c = 0; c < 3; c++) {
            d[c][0] = T.U(c, k, j, i + 1) - T.U(c, k, j, i);
            if (T.NV(k, j + 1, i) > SOL || T.NV(k, j + 1, i + 1) > SOL)
                d[c][1] = (T.U(c, k, j, i + 1) + T.U(c, k, j, i) - T.U(c, k, j - 1, i + 1) - T.U(c, k, j - 1, i)) * 0.5;
            else if (T.NV(k, j - 1, i) > SOL || T.NV(k, j - 1, i + 1) > SOL)
                d[c][1] = (T.U(c, k, j + 1, i + 1) + T.U(c, k, j + 1, i) - T.U(c, k, j, i + 1) - T.U(c, k, j, i)) * 0.5;
            else
                d[c][1] = (T.U(c, k, j + 1, i + 1) + T.U(c, k, j + 1, i) - T.U(c, k, j - 1, i + 1) - T.U(c, k, j - 1, i)) * 0.25;
            if (T.NV(k + 1, j, i) > SOL || T.NV(k + 1, j, i + 1) > SOL)
                d[c][2] = (T.U(c, k, j, i + 1) + T.U(c, k, j, i) - T.U(c, k - 1, j, i + 1) - T.U(c, k - 1, j, i)) * 0.5;
            else if (T.NV(k - 1, j, i) > SOL || T.NV(k - 1, j, i + 1) > SOL)
                d[c][2] = (T.U(c, k + 1, j, i + 1) + T.U(c, k + 1, j, i) - T.U(c, k, j, i + 1) - T.U(c, k, j, i)) * 0.5;
            else
                d[c][2] = (T.U(c, k + 1, j, i + 1) + T.U(c, k + 1, j, i) - T.U(c, k - 1, j, i + 1) - T.U(c, k - 1, j, i)) * 0.25;
        }
        int iL = i - 1, iR = i + 2;
        if (i == 0 || i == mx - 2) {
            if (g.pi && i == mx - 2) iR = mx + 2;
            else if (g.pi && i == 0) iL = -3;
            else iL = i, iR = i + 1;
        } else if (T.NV(k, j, iL) + T.NV(k, j, iR) > 0.1) iL = i, iR = i + 1;
        if (SO) iL = i, iR = i + 1;
        const double uc = ucf;
        for (int c = 0; c < 3; c++) {
            if (SO) dv[c] = -uc * 0.5 * (T.U(c, k, j, i) + T.U(c, k, j, i + 1));
            else dv[c] = -uc * 0.0625 * (-T.U(c, k, j, iL) + 9. * T.U(c, k, j, i) + 9. * T.U(c, k, j, i + 1) - T.U(c, k, j, iR));
        }
        if (T.NV(k, j, i) + T.NV(k, j, i + 1) > 0.1) dv[0] = dv[1] = dv[2] = 0;
        double nut = 0;
        if (a.les) {
            if ((i == 0 && !g.pi) || T.NV(k, j, i) > 0.1) nut = T.NT(k, j, i + 1);
            else if ((i == mx - 2 && !g.pi) || T.NV(k, j, i + 1) > 0.1) nut = T.NT(k, j, i);
            else nut = 0.5 * (T.NT(k, j, i) + T.NT(k, j, i + 1));
        }
        visc_face(cs, et, ze, cs, Mf[9], d, a.nu, nut, a.les, a.laplacian, vs);
    }
}

// convective and viscous flux of the j face of point (i,j,k)
// (momentum.c:930-1211): j in [0,my-2], i,k interior; zero elsewhere
template <bool SO, bool NB = false>
__device__ inline void flux_j(const FaceArgs &a, const TileAcc<SO, NB> &T, int i, int j, int k, const double *Mf, double ucf,
                              double dv[3], double vs[3])
{
    const Grid &g = a.g;
    const int mx = g.mx, my = g.my, mz = g.mz;
    const double SOL = 0.1;  // k-omega.c:19
    (void)SOL; (void)my; (void)mz;
    for (int c = 0; c < 3; c++) dv[c] = vs[c] = 0;
    if (i >= 1 && i < mx - 1 && j < my - 1 && k >= 1 && k < mz - 1) {
        const double cs[3] = {Mf[0], Mf[1], Mf[2]};
        const double et[3] = {Mf[3], Mf[4], Mf[5]};
        const double ze[3] = {Mf[6], Mf[7], Mf[8]};
        double d[3][3];
        // Compute_du_j, k-omega.c:141-200
        for (int c = 0; c < 3; c++) {
            if (T.NV(k, j, i + 1) > SOL || T.NV(k, j + 1, i + 1) > SOL)
                d[c][0] = (T.U(c, k, j + 1, i) + T.U(c, k, j, i) - T.U(c, k, j + 1, i - 1) - T.U(c, k, j, i - 1)) * 0.5;
            else if (T.NV(k, j, i - 1) > SOL || T.NV(k, j + 1, i - 1) > SOL)
                d[c][0] = (T.U(c, k, j + 1, i + 1) + T.U(c, k, j, i + 1) - T.U(c, k, j + 1, i) - T.U(c, k, j, i)) * 0.5;
            else
                d[c][0] = (T.U(c, k, j + 1, i + 1) + T.U(c, k, j, i + 1) - T.U(c, k, j + 1, i - 1) - T.U(c, k, j, i - 1)) * 0.25;
            d[c][1] = T.U(c, k, j + 1, i) - T.U(c, k, j, i);
            if (T.NV(k + 1, j, i) > SOL || T.NV(k + 1, j + 1, i) > SOL)
                d[c][2] = (T.U(c, k, j + 1, i) + T.U(c, k, j, i) - T.U(c, k - 1, j + 1, i) - T.U(c, k - 1, j, i)) * 0.5;
            else if (T.NV(k - 1, j, i) > SOL || T.NV(k - 1, j + 1, i) > SOL)
                d[c][2] = (T.U(c, k + 1, j + 1, i) + T.U(c, k + 1, j, i) - T.U(c, k, j + 1, i) - T.U(c, k, j, i)) * 0.5;
            else
                d[c][2] = (T.U(c, k + 1, j + 1, i) + T.U(c, k + 1, j, i) - T.U(c, k - 1, j + 1, i) - T.U(c, k - 1, j, i)) * 0.25;
        }
        int jL = j - 1, jR = j + 2;
        if (j == 0 || j == my - 2) {
            if (g.pj && j == my - 2) jR = my + 2;
            else if (g.pj && j == 0) jL = -3;
            else jL = j, jR = j + 1;
        } else if (T.NV(k, jL, i) + T.NV(k, jR, i) > 0.1) jL = j, jR = j + 1;
        if (SO) jL = j, jR = j + 1;
        const double uc = ucf;
        for (int c = 0; c < 3; c++) {
            if (SO) dv[c] = -uc * 0.5 * (T.U(c, k, j, i) + T.U(c, k, j + 1, i));
            else dv[c] = -uc * 0.0625 * (-T.U(c, k, jL, i) + 9. * T.U(c, k, j, i) + 9. * T.U(c, k, j + 1, i) - T.U(c, k, jR, i));
        }
        if (T.NV(k, j, i) + T.NV(k, j + 1, i) > 0.1) dv[0] = dv[1] = dv[2] = 0;
        double nut = 0;
        if (a.les) {
            if ((j == 0 && !g.pj) || T.NV(k, j, i) > 0.1) nut = T.NT(k, j + 1, i);
            else if ((j == my - 2 && !g.pj) || T.NV(k, j + 1, i) > 0.1) nut = T.NT(k, j, i);
            else nut = 0.5 * (T.NT(k, j, i) + T.NT(k, j + 1, i));
        }
        visc_face(cs, et, ze, et, Mf[9], d, a.nu, nut, a.les, a.laplacian, vs);
    }
}

// convective and viscous flux of the k face of point (i,j,k)
// (momentum.c:1212-1451): k in [0,mz-2], i,j interior; zero elsewhere
template <bool SO, bool NB = false>
__device__ inline void flux_k(const FaceArgs &a, const TileAcc<SO, NB> &T, int i, int j, int k, const double *Mf, double ucf,
                              double dv[3], double vs[3])
{
    const Grid &g = a.g;
    const int mx = g.mx, my = g.my, mz = g.mz;
    const double SOL = 0.1;  // k-omega.c:19
    (void)SOL; (void)my; (void)mz;
    for (int c = 0; c < 3; c++) dv[c] = vs[c] = 0;
    if (i >= 1 && i < mx - 1 && j >= 1 && j < my - 1 && k < mz - 1) {
        const double cs[3] = {Mf[0], Mf[1], Mf[2]};
        const double et[3] = {Mf[3], Mf[4], Mf[5]};
        const double ze[3] = {Mf[6], Mf[7], Mf[8]};
        double d[3][3];
        // Compute_du_k, k-omega.c:227-288
        for (int c = 0; c < 3; c++) {
            if (T.NV(k, j, i + 1) > SOL || T.NV(k + 1, j, i + 1) > SOL)
                d[c][0] = (T.U(c, k + 1, j, i) + T.U(c, k, j, i) - T.U(c, k + 1, j, i - 1) - T.U(c, k, j, i - 1)) * 0.5;
            else if (T.NV(k, j, i - 1) > SOL || T.NV(k + 1, j, i - 1) > SOL)
                d[c][0] = (T.U(c, k + 1, j, i + 1) + T.U(c, k, j, i + 1) - T.U(c, k + 1, j, i) - T.U(c, k, j, i)) * 0.5;
            else
                d[c][0] = (T.U(c, k + 1, j, i + 1) + T.U(c, k, j, i + 1) - T.U(c, k + 1, j, i - 1) - T.U(c, k, j, i - 1)) * 0.25;
            if (T.NV(k, j + 1, i) > SOL || T.NV(k + 1, j + 1, i) > SOL)
                d[c][1] = (T.U(c, k + 1, j, i) + T.U(c, k, j, i) - T.U(c, k + 1, j - 1, i) - T.U(c, k, j - 1, i)) * 0.5;
            else if (T.NV(k, j - 1, i) > SOL || T.NV(k + 1, j - 1, i) > SOL)
                d[c][1] = (T.U(c, k + 1, j + 1, i) + T.U(c, k, j + 1, i) - T.U(c, k + 1, j, i) - T.U(c, k, j, i)) * 0.5;
            else
                d[c][1] = (T.U(c, k + 1, j + 1, i) + T.U(c, k, j + 1, i) - T.U(c, k + 1, j - 1, i) - T.U(c, k, j - 1, i)) * 0.25;
            d[c][2] = T.U(c, k + 1, j, i) - T.U(c, k, j, i);
        }
        int kL = k - 1, kR = k + 2;
        if (k == 0 || k == mz - 2) {
            if (g.pk && k == mz - 2) kR = mz + 2;
            else if (g.pk && k == 0) kL = -3;
            else kL = k, kR = k + 1;
        } else if (T.NV(kL, j, i) + T.NV(kR, j, i) > 0.1) kL = k, kR = k + 1;
        if (SO) kL = k, kR = k + 1;
        const double uc = ucf;
        for (int c = 0; c < 3; c++) {
            if (SO) dv[c] = -uc * 0.5 * (T.U(c, k, j, i) + T.U(c, k + 1, j, i));
            else dv[c] = -uc * 0.0625 * (-T.U(c, kL, j, i) + 9. * T.U(c, k, j, i) + 9. * T.U(c, k + 1, j, i) - T.U(c, kR, j, i));
        }
        if (T.NV(k, j, i) + T.NV(k + 1, j, i) > 0.1) dv[0] = dv[1] = dv[2] = 0;
        double nut = 0;
        if (a.les) {
            if ((k == 0 && !g.pk) || T.NV(k, j, i) > 0.1) nut = T.NT(k + 1, j, i);
            else if ((k == mz - 2 && !g.pk) || T.NV(k + 1, j, i) > 0.1) nut = T.NT(k, j, i);
            else nut = 0.5 * (T.NT(k, j, i) + T.NT(k + 1, j, i));
        }
        visc_face(cs, et, ze, ze, Mf[9], d, a.nu, nut, a.les, a.laplacian, vs);
    }
}

// the 18 face fluxes of point (i,j,k) (momentum.c:668-1451); SO selects the
// second-order convection at compile time (a.second_order is then 1).  With
// DER the face metrics are rebuilt from the cell metrics: `cur` holds the
// point's own cell metrics and `nk` those of cell k+1.
template <bool SO, bool DER>
__device__ inline void face_point(const FaceArgs &a, const TileAcc<SO> &T, int i, int j, int k, const double *cur,
                                  const double *nk)
{
    const Grid &g = a.g;
    const long long q = ((long long)k * g.my + j) * g.mx + i;
    const long long n = g.n;
    const int mx = g.mx, my = g.my, mz = g.mz;
    const double SOL = 0.1;  // k-omega.c:19
    // the point's three face-metric families and fluxes, loaded together up
    // front (q is a valid point for every face, used or not) so that the 33
    // HBM loads are in flight at once
    double M[3][10], UC[3];
#pragma unroll
    for (int f = 0; f < 3; f++) UC[f] = a.ucont[3 * q + f];
    if (DER) {
        // neighbour cells i+1 and j+1 (the point itself where that face is
        // not computed, which keeps the loads inside the array)
        double ni[10], nj[10];
        load_metrics(a.cell, i < mx - 1 ? q + 1 : q, ni);
        load_metrics(a.cell, j < my - 1 ? q + mx : q, nj);
        face_mean(cur, ni, M[0]);
        face_mean(cur, nj, M[1]);
        face_mean(cur, nk, M[2]);
    } else {
        load_metrics(a.fi, q, M[0]);
        load_metrics(a.fj, q, M[1]);
        load_metrics(a.fk, q, M[2]);
    }

    double dv[3], vs[3];
    flux_i<SO>(a, T, i, j, k, M[0], UC[0], dv, vs);
    for (int c = 0; c < 3; c++) { a.out[(0 + c) * n + q] = dv[c]; a.out[(9 + c) * n + q] = vs[c]; }
    flux_j<SO>(a, T, i, j, k, M[1], UC[1], dv, vs);
    for (int c = 0; c < 3; c++) { a.out[(3 + c) * n + q] = dv[c]; a.out[(12 + c) * n + q] = vs[c]; }
    flux_k<SO>(a, T, i, j, k, M[2], UC[2], dv, vs);
    for (int c = 0; c < 3; c++) { a.out[(6 + c) * n + q] = dv[c]; a.out[(15 + c) * n + q] = vs[c]; }
}

// One block owns an FT_I x FT_J column of points and marches FT_K planes in
// k, keeping a rolling 3-plane LDS tile of ucat / nvert / nu_t.
template <bool SO, bool DER>
__global__ void __launch_bounds__(FT_I * FT_J) k_face_fluxes(FaceArgs a)
{
    using D = TileDims<SO>;
    __shared__ double sx[3 * D::HP], sy[3 * D::HP], sz[3 * D::HP], st[3 * D::HP];
    __shared__ float sn[3 * D::HP];
    const Grid &g = a.g;
    const int nbi = (g.mx + FT_I - 1) / FT_I, nbj = (g.my + FT_J - 1) / FT_J;
    int b = blockIdx.x;
    const int bi = b % nbi;
    b /= nbi;
    const int bj = b % nbj, bk = b / nbj;
    const int i0 = bi * FT_I, j0 = bj * FT_J, k0 = bk * FT_K;
    const int k1 = k0 + FT_K < g.mz ? k0 + FT_K : g.mz;
    const int tx = threadIdx.x % FT_I, ty = threadIdx.x / FT_I;
    const bool les = a.les != 0;
    auto load_plane = [&](int kk, int slot) {
        const int wk = wrapi(kk, g.mz, g.pk);
        for (int p = threadIdx.x; p < D::HP; p += FT_I * FT_J) {
            const int di = p % D::HI, dj = p / D::HI;
            const int wi = wrapi(i0 - 1 + di, g.mx, g.pi), wj = wrapi(j0 - 1 + dj, g.my, g.pj);
            double vx = 0, vy = 0, vz = 0, vt = 0;
            float vn = 0;
            if (wi >= 0 && wi < g.mx && wj >= 0 && wj < g.my && wk >= 0 && wk < g.mz) {
                const long long q = ((long long)wk * g.my + wj) * g.mx + wi;
                vx = a.ucat[3 * q]; vy = a.ucat[3 * q + 1]; vz = a.ucat[3 * q + 2];
                vn = (float)a.nvert[q];
                if (les) vt = a.nut[q];
            }
            const int l = slot * D::HP + p;
            sx[l] = vx; sy[l] = vy; sz[l] = vz; sn[l] = vn; st[l] = vt;
        }
    };
    load_plane(k0 - 1, 0);
    load_plane(k0, 1);
    int slot0 = 0;
    // with DER the cell metrics of k+1 loaded for one plane are the point's
    // own on the next
    double cur[DER ? 10 : 1], nk[DER ? 10 : 1];
    bool have = false;
    for (int k = k0; k < k1; k++) {
        int s2 = slot0 + 2;
        if (s2 >= 3) s2 -= 3;
        load_plane(k + 1, s2);
        __syncthreads();
        const int i = i0 + tx, j = j0 + ty;
        if (i < g.mx && j < g.my) {
            if (DER) {
                const long long q = ((long long)k * g.my + j) * g.mx + i;
                if (!have) load_metrics(a.cell, q, cur);
                load_metrics(a.cell, k < g.mz - 1 ? q + g.plane : q, nk);
            }
            TileAcc<SO> T{&a, sx, sy, sz, st, sn, i0, j0, k, slot0};
            face_point<SO, DER>(a, T, i, j, k, cur, nk);
            if (DER) {
#pragma unroll
                for (int c = 0; c < 10; c++) cur[c] = nk[c];
                have = true;
            }
        }
        __syncthreads();
        slot0 = slot0 + 1 < 3 ? slot0 + 1 : 0;
    }
}

// The face pass fused with the flux divergence (momentum.c:1548-1678) for
// second-order convection with the face metrics rebuilt from the cell
// metrics.  A block owns an FT_I x FT_J column and marches in k as
// k_face_fluxes does; on each plane the i- and j-face fluxes of its points go
// to LDS together with one extra column (the i face at i0-1) and one extra row
// (the j face at j0-1), and the k-face flux of the plane below stays in
// registers, so Fp = (div_i(i) - div_i(i-1) + ... ) + (visc ...) is formed in
// the reference's order without the 18 flux planes ever reaching HBM.  What
// is stored per cell is not Fp but its three projections on the cell's own
// contravariant directions, csi.Fp, eta.Fp and zet.Fp: the projection pass
// (momentum.c:1715-1841) uses Fp only through these products (of a cell's
// Fp with that cell's metrics), evaluated here with the same operands in the
// same order.  A block starting above k = 0 first computes the k face of
// plane k0-1 only.
// Periodic ghost cells of the face fluxes are copies of the interior faces
// they stand for (momentum.c:1504-1546); with the periodic ghosts of ucat,
// nu_t, nvert, ucont and the metrics holding their periodic copies (as
// Contra2Cart, the eddy viscosity, Formfunction_2's own Ucont fix and
// FormMetrics leave them) the faces at index 0 computed here are bitwise
// those copies, so that fix is not applied.  The projections are written for
// every point of the block (zero at ghosts); the periodic ghosts are filled
// afterwards.
constexpr int FX_J = FT_J + 1;
static_assert(FT_I == 32, "a row of the tile is half a wavefront (the i-face exchange shuffles within it)");
constexpr int HX_T0 = 64;  // first thread of the second wavefront
static_assert(FT_I * FT_J >= HX_T0 + FT_J && FT_J * 2 <= 64, "halo lanes of the i faces");
inline unsigned face_fp_blocks(const Grid &g)
{
    return (unsigned)(((g.mx - 2 + FT_I - 1) / FT_I) * ((g.my - 2 + FT_J - 1) / FT_J) * ((g.mz - 2 + FT_K - 1) / FT_K));
}
template <bool NB>
__global__ void __launch_bounds__(FT_I * FT_J) __attribute__((amdgpu_waves_per_eu(2))) k_face_fp(FaceArgs a, double *__restrict__ fp)
{
    using D = TileDims<true>;
    __shared__ double sx[3 * D::HP], sy[3 * D::HP], sz[3 * D::HP], st[3 * D::HP];
    __shared__ float sn[3 * D::HP];
    __shared__ double fxj[6][FX_J][FT_I];  // j-face div xyz, visc xyz; row 0 = face j0-1
    __shared__ double hx[6][FT_J];         // i-face div xyz, visc xyz at i0-1 of each row
    const Grid &g = a.g;
    // The tiles cover the interior points (1 .. m-2 on each axis; face_fp_blocks
    // blocks): the faces at index 0 are the halo faces of the first tiles, and
    // the ghost points' projections are not written (zero at the non-periodic
    // ghosts since allocation, copies at the periodic ones by the periodic fix
    // that follows).  Tiles are taken in XCD order, so the tiles that share a
    // halo run on one XCD's L2.
    const int nbi = (g.mx - 2 + FT_I - 1) / FT_I, nbj = (g.my - 2 + FT_J - 1) / FT_J;
    int b = (int)xcd_block1();
    const int bi = b % nbi;
    b /= nbi;
    const int bj = b % nbj, bk = b / nbj;
    const int i0 = 1 + bi * FT_I, j0 = 1 + bj * FT_J, k0 = 1 + bk * FT_K;
    const int k1 = k0 + FT_K < g.mz - 1 ? k0 + FT_K : g.mz - 1;
    const int tx = threadIdx.x % FT_I, ty = threadIdx.x / FT_I;
    const int i = i0 + tx, j = j0 + ty;
    const bool inb = i < g.mx && j < g.my;
    const long long n = g.n;
    const bool les = a.les != 0;
    // A tile plane is fetched into registers one plane ahead and written to
    // LDS at the top of the next iteration, so the global loads of plane k+2
    // are in flight while plane k is computed.  Each thread owns tile points
    // tid and tid + FT_I*FT_J.
    constexpr int TPP = (D::HP + FT_I * FT_J - 1) / (FT_I * FT_J);  // tile points per thread
    double rv[TPP][4];
    float rn[TPP];
    auto fetch_plane = [&](int kk) {
        const int wk = wrapi(kk, g.mz, g.pk);
#pragma unroll
        for (int r = 0; r < TPP; r++) {
            const int p = threadIdx.x + r * FT_I * FT_J;
            rv[r][0] = rv[r][1] = rv[r][2] = rv[r][3] = 0;
            rn[r] = 0;
            if (p < D::HP) {
                const int di = p % D::HI, dj = p / D::HI;
                const int wi = wrapi(i0 - 1 + di, g.mx, g.pi), wj = wrapi(j0 - 1 + dj, g.my, g.pj);
                if (wi >= 0 && wi < g.mx && wj >= 0 && wj < g.my && wk >= 0 && wk < g.mz) {
                    const long long q = ((long long)wk * g.my + wj) * g.mx + wi;
                    rv[r][0] = a.ucat[3 * q]; rv[r][1] = a.ucat[3 * q + 1]; rv[r][2] = a.ucat[3 * q + 2];
                    if (!NB) rn[r] = (float)a.nvert[q];
                    if (les) rv[r][3] = a.nut[q];
                }
            }
        }
    };
    auto store_plane = [&](int slot) {
#pragma unroll
        for (int r = 0; r < TPP; r++) {
            const int p = threadIdx.x + r * FT_I * FT_J;
            if (p < D::HP) {
                const int l = slot * D::HP + p;
                sx[l] = rv[r][0]; sy[l] = rv[r][1]; sz[l] = rv[r][2]; st[l] = rv[r][3];
                if (!NB) sn[l] = rn[r];
            }
        }
    };
    const int kstart = k0 - 1;
    fetch_plane(kstart - 1);
    store_plane(0);
    fetch_plane(kstart);
    store_plane(1);
    fetch_plane(kstart + 1);
    int slot0 = 0;
    double cur[10], nk[10], kprev[6];
    for (int c = 0; c < 6; c++) kprev[c] = 0;
    bool have = false;
    for (int k = kstart; k < k1; k++) {
        int s2 = slot0 + 2;
        if (s2 >= 3) s2 -= 3;
        store_plane(s2);  // plane k+1
        __syncthreads();
        if (k + 1 < k1) fetch_plane(k + 2);
        TileAcc<true, NB> T{&a, sx, sy, sz, st, sn, i0, j0, k, slot0};
        const bool full = k >= k0;  // the plane below the block: its k face only
        const long long q = ((long long)k * g.my + j) * g.mx + i;
        if (inb && !have) load_metrics(a.cell, q, cur);
        double di[6] = {0, 0, 0, 0, 0, 0};  // i-face differences, div xyz and visc xyz
        if (full) {
            // i faces: this point's face i+1/2 and, from the lane before it
            // in the row (a row of FT_I points is FT_I consecutive lanes of
            // one wavefront), the face i-1/2; the row's first lane takes the
            // face i0-1/2 from LDS after the barrier.
            double f6[6] = {0, 0, 0, 0, 0, 0};
            if (inb) {
                double nb_[10], M[10];
                load_metrics(a.cell, i < g.mx - 1 ? q + 1 : q, nb_);
                face_mean(cur, nb_, M);
                flux_i<true, NB>(a, T, i, j, k, M, a.ucont[3 * q + 0], f6, f6 + 3);
            }
            double fl[6];
            for (int c = 0; c < 6; c++) fl[c] = __shfl_up(f6[c], 1, FT_I);
            for (int c = 0; c < 6; c++) di[c] = tx == 0 ? f6[c] : f6[c] - fl[c];
            for (int c = 0; c < 6; c++) f6[c] = 0;
            if (inb) {
                double nb_[10], M[10];
                load_metrics(a.cell, j < g.my - 1 ? q + g.mx : q, nb_);
                face_mean(cur, nb_, M);
                flux_j<true, NB>(a, T, i, j, k, M, a.ucont[3 * q + 1], f6, f6 + 3);
            }
            for (int c = 0; c < 6; c++) fxj[c][ty + 1][tx] = f6[c];
            if (ty == 0) {  // j face at j0-1 of this column
                for (int c = 0; c < 6; c++) f6[c] = 0;
                if (j0 >= 1 && i < g.mx) {
                    const long long p = q - g.mx;
                    double cp[10], M[10];
                    load_metrics(a.cell, p, cp);
                    face_mean(cp, cur, M);
                    flux_j<true, NB>(a, T, i, j0 - 1, k, M, a.ucont[3 * p + 1], f6, f6 + 3);
                }
                for (int c = 0; c < 6; c++) fxj[c][0][tx] = f6[c];
            }
            // the face i0-1/2 of each row, on lanes of the second wavefront
            // (the first one computes the faces j0-1/2), so that no
            // wavefront evaluates more than one halo face
            const int r = (int)threadIdx.x - HX_T0;
            if (r >= 0 && r < FT_J) {
                const int jr = j0 + r;
                for (int c = 0; c < 6; c++) f6[c] = 0;
                if (i0 >= 1 && jr < g.my) {
                    const long long p = ((long long)k * g.my + jr) * g.mx + (i0 - 1);
                    double cp[10], cn[10], M[10];
                    load_metrics(a.cell, p, cp);
                    load_metrics(a.cell, p + 1, cn);
                    face_mean(cp, cn, M);
                    flux_i<true, NB>(a, T, i0 - 1, jr, k, M, a.ucont[3 * p + 0], f6, f6 + 3);
                }
                for (int c = 0; c < 6; c++) hx[c][r] = f6[c];
            }
        }
        double kf[6] = {0, 0, 0, 0, 0, 0};
        if (inb) {
            load_metrics(a.cell, k < g.mz - 1 ? q + g.plane : q, nk);
            double M[10];
            face_mean(cur, nk, M);
            flux_k<true, NB>(a, T, i, j, k, M, a.ucont[3 * q + 2], kf, kf + 3);
        }
        __syncthreads();
        if (full && inb) {
            double f[3] = {0, 0, 0};
            double pr[3] = {0, 0, 0};
            if (tx == 0)
                for (int c = 0; c < 6; c++) di[c] = di[c] - hx[c][ty];
            if (i >= 1 && i < g.mx - 1 && j >= 1 && j < g.my - 1 && k >= 1 && k < g.mz - 1) {
                for (int c = 0; c < 3; c++) {
                    f[c] = di[c] + fxj[c][ty + 1][tx] - fxj[c][ty][tx] + kf[c] - kprev[c];
                    f[c] += (di[3 + c] + fxj[3 + c][ty + 1][tx] - fxj[3 + c][ty][tx] + kf[3 + c] - kprev[3 + c]) / 1.0;
                }
                // csi.Fp, eta.Fp, zet.Fp of this cell
                for (int r = 0; r < 3; r++) pr[r] = cur[3 * r] * f[0] + cur[3 * r + 1] * f[1] + cur[3 * r + 2] * f[2];
            }
            for (int r = 0; r < 3; r++) fp[r * n + q] = pr[r];
        }
        for (int c = 0; c < 6; c++) kprev[c] = kf[c];
        if (inb) {
            for (int c = 0; c < 10; c++) cur[c] = nk[c];
            have = true;
        }
        // No barrier here: the next store_plane overwrites the slot of plane
        // k-1, which only the flux evaluations read (all done before the
        // barrier above), and fxj / hx are next written after the next
        // iteration's first barrier, which every wavefront reaches only once
        // it has finished reading them here.
        slot0 = slot0 + 1 < 3 ? slot0 + 1 : 0;
    }
}

// periodic fix of the 18 flux planes (momentum.c:1504-1546)
__global__ void k_periodic_soa(Grid g, double *__restrict__ f, int nc, int dir)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    int i, j, k;
    if (dir == 0) {
        if (t >= 2LL * g.my * g.mz) return;
        int side = t & 1; long long r = t >> 1;
        j = r % g.my; k = r / g.my; i = side ? g.mx - 1 : 0;
    } else if (dir == 1) {
        if (t >= 2LL * g.mx * g.mz) return;
        i = t % g.mx; long long r = t / g.mx;
        int side = r & 1; k = r >> 1; j = side ? g.my - 1 : 0;
    } else {
        if (t >= 2LL * g.mx * g.my) return;
        i = t % g.mx; long long r = t / g.mx;
        j = r % g.my; int side = (int)(r / g.my); k = side ? g.mz - 1 : 0;
    }
    int a, b, c, fi, fj, fk;
    if (!periodic_src(g, k, j, i, c, b, a, fi, fj, fk)) return;
    long long d = gidx(g, k, j, i), s = gidx(g, c, b, a);
    for (int q = 0; q < nc; q++) f[q * g.n + d] = f[q * g.n + s];
}

struct CellArgs {
    Grid g;
    const double *face;  // 18 planes
    const double *csi, *eta, *zet, *iaj, *jaj, *kaj, *nvert;
    double scale;
    int second_order;
    double *rhs;          // Rhs (accumulated) when pucont == nullptr
    const double *dp;     // RK mode: dP
    const double *pucont; // RK mode: pUcont
    double alpha;         // RK mode: alfa * dt * st
    double *ucont_out;    // RK mode: Ucont
    // CN mode (cn_r != nullptr; k_cell_rhs_pr only): FormFunction_SNES's
    // residual r = (0 - u/dt + u_o/dt) + F + 0.5 RHS_o - dP (+ F_eul)
    const double *cn_u, *cn_uo, *cn_rhso, *cn_feul;
    double cn_rdt;
    double *cn_r;
    const double *cn_rbase;  // set: cn_r = cn_da cn_rbase + cn_db R
    double cn_da, cn_db;
};

__device__ inline double Fv(const CellArgs &a, int plane, int k, int j, int i)
{
    return a.face[plane * a.g.n + gidx(a.g, k, j, i)];
}

// Fp of an interior cell, momentum.c:1548-1678; SO is a.second_order
template <bool SO>
__device__ void fp_interior(const CellArgs &a, int k, int j, int i, double f[3])
{
    const Grid &g = a.g;
    const int mx = g.mx, my = g.my, mz = g.mz;
    double dv[3];
    for (int c = 0; c < 3; c++)
        dv[c] = (Fv(a, c, k, j, i) - Fv(a, c, k, j, i - 1) + Fv(a, 3 + c, k, j, i) - Fv(a, 3 + c, k, j - 1, i) +
                 Fv(a, 6 + c, k, j, i) - Fv(a, 6 + c, k - 1, j, i));
    if (SO) {
        for (int c = 0; c < 3; c++) f[c] = dv[c];
    } else {
        int iR = i + 1, iL = i - 2, jR = j + 1, jL = j - 2, kR = k + 1, kL = k - 2;
        double di = 3., dj = 3., dk = 3.;
        if (i == 1) { if (g.pi) iL = -3; else iR = i, iL = i - 1, di = 1.; }
        else if (i == 2 || i == mx - 3) { if (!g.pi) iR = i, iL = i - 1, di = 1.; }
        else if (i == mx - 2) { if (g.pi) iR = mx + 1; else iR = i, iL = i - 1, di = 1.; }
        if (j == 1) { if (g.pj) jL = -3; else jR = j, jL = j - 1, dj = 1.; }
        else if (j == 2 || j == my - 3) { if (!g.pj) jR = j, jL = j - 1, dj = 1.; }
        else if (j == my - 2) { if (g.pj) jR = my + 1; else jR = j, jL = j - 1, dj = 1.; }
        if (k == 1) { if (g.pk) kL = -3; else kR = k, kL = k - 1, dk = 1.; }
        else if (k == 2 || k == mz - 3) { if (!g.pk) kR = k, kL = k - 1, dk = 1.; }
        else if (k == mz - 2) { if (g.pk) kR = mz + 1; else kR = k, kL = k - 1, dk = 1.; }
        const double *nv = a.nvert;
        if (nv[gidx(g, k, j, i - 1)] + nv[gidx(g, k, j, i)] + nv[gidx(g, k, j, i + 1)] > 0.1) iR = i, iL = i - 1, di = 1.;
        if (nv[gidx(g, k, j - 1, i)] + nv[gidx(g, k, j, i)] + nv[gidx(g, k, j + 1, i)] > 0.1) jR = j, jL = j - 1, dj = 1.;
        if (nv[gidx(g, k - 1, j, i)] + nv[gidx(g, k, j, i)] + nv[gidx(g, k + 1, j, i)] > 0.1) kR = k, kL = k - 1, dk = 1.;
        for (int c = 0; c < 3; c++) {
            double d4 = 0;
            d4 += (Fv(a, c, k, j, iR) - Fv(a, c, k, j, iL)) * (1. / di);
            d4 += (Fv(a, 3 + c, k, jR, i) - Fv(a, 3 + c, k, jL, i)) * (1. / dj);
            d4 += (Fv(a, 6 + c, kR, j, i) - Fv(a, 6 + c, kL, j, i)) * (1. / dk);
            f[c] = 9. / 8. * dv[c] + -1. / 8. * d4;
        }
    }
    for (int c = 0; c < 3; c++)
        f[c] += (Fv(a, 9 + c, k, j, i) - Fv(a, 9 + c, k, j, i - 1) + Fv(a, 12 + c, k, j, i) - Fv(a, 12 + c, k, j - 1, i) +
                 Fv(a, 15 + c, k, j, i) - Fv(a, 15 + c, k - 1, j, i)) / 1.0;
}

// Fp as the projection loop reads it: interior cells directly, periodic ghost
// cells through the periodic fix (momentum.c:1685-1713), other ghosts 0.
template <bool SO>
__device__ inline void fp_at(const CellArgs &a, int k, int j, int i, double f[3])
{
    const Grid &g = a.g;
    int c_, b_, a_, fi, fj, fk;
    if (periodic_src(g, k, j, i, c_, b_, a_, fi, fj, fk)) { k = c_; j = b_; i = a_; }
    if (i < 1 || i > g.mx - 2 || j < 1 || j > g.my - 2 || k < 1 || k > g.mz - 2) {
        f[0] = f[1] = f[2] = 0;
        return;
    }
    fp_interior<SO>(a, k, j, i, f);
}

// Fp of every point as the projection reads it (fp_at), written once to a
// 3-plane scratch.  A block owns an FT_I x FT_J column and marches FT_K planes
// in k; on the second-order path a thread keeps its k-face fluxes in
// registers, where they become the k-1 face of the next plane.
template <bool SO>
__global__ void __launch_bounds__(FT_I * FT_J) k_cell_fp(CellArgs a, double *__restrict__ fp)
{
    const Grid &g = a.g;
    const int nbi = (g.mx + FT_I - 1) / FT_I, nbj = (g.my + FT_J - 1) / FT_J;
    int b = blockIdx.x;
    const int bi = b % nbi;
    b /= nbi;
    const int bj = b % nbj, bk = b / nbj;
    const int i = bi * FT_I + threadIdx.x % FT_I, j = bj * FT_J + threadIdx.x / FT_I, k0 = bk * FT_K;
    if (i >= g.mx || j >= g.my) return;
    const int k1 = k0 + FT_K < g.mz ? k0 + FT_K : g.mz;
    const long long n = g.n;
    double km1[6];
    bool have = false;
    for (int k = k0; k < k1; k++) {
        const long long q = ((long long)k * g.my + j) * g.mx + i;
        double f[3];
        const bool interior = i >= 1 && i < g.mx - 1 && j >= 1 && j < g.my - 1 && k >= 1 && k < g.mz - 1;
        if (interior && SO) {
            const long long qi = q - 1, qj = q - g.mx;
            const double *__restrict__ F = a.face;
            // all loads first (18 own planes, 6 at i-1, 6 at j-1), then the
            // arithmetic, so the loads are in flight together
            double own[18], im[6], jm[6];
#pragma unroll
            for (int p = 0; p < 18; p++) own[p] = F[p * n + q];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                im[c] = F[c * n + qi];
                im[3 + c] = F[(9 + c) * n + qi];
                jm[c] = F[(3 + c) * n + qj];
                jm[3 + c] = F[(12 + c) * n + qj];
            }
            if (!have) {
                const long long qk = q - g.plane;
#pragma unroll
                for (int c = 0; c < 3; c++) { km1[c] = F[(6 + c) * n + qk]; km1[3 + c] = F[(15 + c) * n + qk]; }
            }
            // same operation order as fp_interior
#pragma unroll
            for (int c = 0; c < 3; c++) {
                f[c] = own[c] - im[c] + own[3 + c] - jm[c] + own[6 + c] - km1[c];
                f[c] += (own[9 + c] - im[3 + c] + own[12 + c] - jm[3 + c] + own[15 + c] - km1[3 + c]) / 1.0;
            }
#pragma unroll
            for (int c = 0; c < 3; c++) { km1[c] = own[6 + c]; km1[3 + c] = own[15 + c]; }
            have = true;
        } else {
            fp_at<SO>(a, k, j, i, f);
            have = false;
        }
        for (int c = 0; c < 3; c++) fp[c * n + q] = f[c];
    }
}

// Contravariant projection of Fp and boundary zeroing (momentum.c:1715-1938),
// marching in k with the k+1 neighbour carried in registers.
__global__ void __launch_bounds__(FT_I * FT_J) k_cell_rhs(CellArgs a, const double *__restrict__ fp)
{
    const Grid &g = a.g;
    const int nbi = (g.mx + FT_I - 1) / FT_I, nbj = (g.my + FT_J - 1) / FT_J;
    int b = blockIdx.x;
    const int bi = b % nbi;
    b /= nbi;
    const int bj = b % nbj, bk = b / nbj;
    const int i = bi * FT_I + threadIdx.x % FT_I, j = bj * FT_J + threadIdx.x / FT_I, k0 = bk * FT_K;
    if (i >= g.mx || j >= g.my) return;
    const int k1 = k0 + FT_K < g.mz ? k0 + FT_K : g.mz;
    const long long n = g.n;
    const int mx = g.mx, my = g.my, mz = g.mz;
    double fk_[3], z1[3], nvk = 0;
    bool have = false;
    for (int k = k0; k < k1; k++) {
        const long long q = ((long long)k * g.my + j) * g.mx + i;
        double r[3];
        const bool interior = i >= 1 && i < mx - 1 && j >= 1 && j < my - 1 && k >= 1 && k < mz - 1;
        bool zero[3] = {true, true, true};
        if (interior) {
            const long long qi = q + 1, qj = q + mx, qk = gidx(g, k + 1, j, i);
            double f0[3], fi_[3], fj_[3];
            for (int c = 0; c < 3; c++) {
                f0[c] = have ? fk_[c] : fp[c * n + q];
                fi_[c] = fp[c * n + qi];
                fj_[c] = fp[c * n + qj];
            }
            const double nv0 = have ? nvk : a.nvert[q];
            double fkn[3], zkn[3];
            for (int c = 0; c < 3; c++) { fkn[c] = fp[c * n + qk]; zkn[c] = a.zet[3 * qk + c]; }
            const double nvkn = a.nvert[qk];
            const double *c0 = a.csi + 3 * q, *c1 = a.csi + 3 * qi;
            const double *e0 = a.eta + 3 * q, *e1 = a.eta + 3 * qj;
            const double *z0 = a.zet + 3 * q;
            r[0] = a.scale * (0.5 * (c0[0] * f0[0] + c0[1] * f0[1] + c0[2] * f0[2]) +
                              0.5 * (c1[0] * fi_[0] + c1[1] * fi_[1] + c1[2] * fi_[2])) * a.iaj[q];
            r[1] = a.scale * (0.5 * (e0[0] * f0[0] + e0[1] * f0[1] + e0[2] * f0[2]) +
                              0.5 * (e1[0] * fj_[0] + e1[1] * fj_[1] + e1[2] * fj_[2])) * a.jaj[q];
            r[2] = a.scale * (0.5 * (z0[0] * f0[0] + z0[1] * f0[1] + z0[2] * f0[2]) +
                              0.5 * (zkn[0] * fkn[0] + zkn[1] * fkn[1] + zkn[2] * fkn[2])) * a.kaj[q];
            const double *nv = a.nvert;
            zero[0] = nv0 + nv[qi] > 0.1 || (!g.pi && i == mx - 2);
            zero[1] = nv0 + nv[qj] > 0.1 || (!g.pj && j == my - 2);
            zero[2] = nv0 + nvkn > 0.1 || (!g.pk && k == mz - 2);
            for (int c = 0; c < 3; c++) { fk_[c] = fkn[c]; z1[c] = zkn[c]; }
            nvk = nvkn;
            have = qk == q + g.plane;  // the carried values belong to the next plane
        } else {
            have = false;
        }
        if (a.pucont) {
            // Rhs = 0 + F; Rhs += -1*dP; Ucont = alpha*Rhs + pUcont  (rhs.c:1647-1658).
            // Every input is loaded before the three stores, which then leave
            // together (a store between loads wrote partial lines to HBM).
            const double *__restrict__ dp = a.dp + 3 * q;
            const double *__restrict__ pu = a.pucont + 3 * q;
            double d[3], u[3], o[3];
#pragma unroll
            for (int c = 0; c < 3; c++) { d[c] = dp[c]; u[c] = pu[c]; }
#pragma unroll
            for (int c = 0; c < 3; c++) {
                double rv = zero[c] ? 0.0 : (0.0 + r[c]);
                rv += -1 * d[c];
                o[c] = a.alpha * rv + u[c];
            }
            double *__restrict__ out = a.ucont_out + 3 * q;
#pragma unroll
            for (int c = 0; c < 3; c++) out[c] = o[c];
        } else if (a.cn_r) {
            // Rhs = 0; Rhs += (-1/dt) u; Rhs += (1/dt) u_o; Rhs += 0.5 F (the
            // cell's r); R = Rhs + 0.5 RHS_o - dP (+ F_eul), in the separate
            // passes' operation order (momentum.c:2237-2336)
            const double *__restrict__ u = a.cn_u + 3 * q;
            const double *__restrict__ uo = a.cn_uo + 3 * q;
            const double *__restrict__ ro = a.cn_rhso + 3 * q;
            const double *__restrict__ dp = a.dp + 3 * q;
            double uu[3], uoo[3], roo[3], d[3], fe[3] = {0, 0, 0}, o[3];
#pragma unroll
            for (int c = 0; c < 3; c++) { uu[c] = u[c]; uoo[c] = uo[c]; roo[c] = ro[c]; d[c] = dp[c]; }
            if (a.cn_feul)
#pragma unroll
                for (int c = 0; c < 3; c++) fe[c] = a.cn_feul[3 * q + c];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double s0 = 0.0 + (-a.cn_rdt) * uu[c];
                const double base = s0 + a.cn_rdt * uoo[c];
                const double rv = zero[c] ? 0.0 : base + r[c];
                double s = rv + 0.5 * roo[c];
                s = s + (-1.0) * d[c];
                if (a.cn_feul) s = s + fe[c];
                o[c] = s;
            }
            if (a.cn_rbase)
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const double rb = a.cn_rbase[3 * q + c];
                    o[c] = a.cn_da * rb + a.cn_db * o[c];
                }
            double *__restrict__ out = a.cn_r + 3 * q;
#pragma unroll
            for (int c = 0; c < 3; c++) out[c] = o[c];
        } else {
            for (int c = 0; c < 3; c++) a.rhs[3 * q + c] = zero[c] ? 0.0 : a.rhs[3 * q + c] + r[c];
        }
    }
    (void)z1;
}

// k_cell_rhs on the projections written by k_face_fp: pr[0..2] planes hold
// csi.Fp, eta.Fp, zet.Fp of every cell, so the contravariant projection
// reads no cell metrics (the same products, hence the same bits).
// The cell pass needs no LDS tile: a block takes CR_T consecutive points of
// a k plane (the plane flattened, so 6 KB of each AoS array per block and no
// block covers only the last few points of a row) and marches FT_K planes.
constexpr int CR_T = 256;
inline unsigned cell_rhs_blocks(const Grid &g)
{
    return (unsigned)(((g.plane + CR_T - 1) / CR_T) * ((g.mz + FT_K - 1) / FT_K));
}
__global__ void __launch_bounds__(CR_T) k_cell_rhs_pr(CellArgs a, const double *__restrict__ pr)
{
    const Grid &g = a.g;
    const long long nbp = (g.plane + CR_T - 1) / CR_T;
    const long long bp = blockIdx.x % nbp, bk = blockIdx.x / nbp;
    const long long pp = bp * CR_T + threadIdx.x;
    if (pp >= g.plane) return;
    const int i = (int)(pp % g.mx), j = (int)(pp / g.mx), k0 = (int)bk * FT_K;
    const int k1 = k0 + FT_K < g.mz ? k0 + FT_K : g.mz;
    const long long n = g.n;
    const int mx = g.mx, my = g.my, mz = g.mz;
    double pk_next = 0, nvk = 0;
    bool have = false;
    for (int k = k0; k < k1; k++) {
        const long long q = ((long long)k * g.my + j) * g.mx + i;
        double r[3];
        const bool interior = i >= 1 && i < mx - 1 && j >= 1 && j < my - 1 && k >= 1 && k < mz - 1;
        bool zero[3] = {true, true, true};
        if (interior) {
            const long long qi = q + 1, qj = q + mx, qk = gidx(g, k + 1, j, i);
            const double pi0 = pr[q], pi1 = pr[qi];
            const double pj0 = pr[n + q], pj1 = pr[n + qj];
            const double pk0 = have ? pk_next : pr[2 * n + q], pk1 = pr[2 * n + qk];
            // a.nvert == nullptr: no point is blanked (every nvert 0)
            const double *nv = a.nvert;
            const double nv0 = have ? nvk : (nv ? nv[q] : 0.0);
            const double nvkn = nv ? nv[qk] : 0.0;
            r[0] = a.scale * (0.5 * pi0 + 0.5 * pi1) * a.iaj[q];
            r[1] = a.scale * (0.5 * pj0 + 0.5 * pj1) * a.jaj[q];
            r[2] = a.scale * (0.5 * pk0 + 0.5 * pk1) * a.kaj[q];
            zero[0] = nv0 + (nv ? nv[qi] : 0.0) > 0.1 || (!g.pi && i == mx - 2);
            zero[1] = nv0 + (nv ? nv[qj] : 0.0) > 0.1 || (!g.pj && j == my - 2);
            zero[2] = nv0 + nvkn > 0.1 || (!g.pk && k == mz - 2);
            nvk = nvkn;
            pk_next = pk1;
            have = qk == q + g.plane;  // the carried values belong to the next plane
        } else {
            have = false;
        }
        if (a.pucont) {
            // Rhs = 0 + F; Rhs += -1*dP; Ucont = alpha*Rhs + pUcont  (rhs.c:1647-1658).
            // Every input is loaded before the three stores, which then leave
            // together (a store between loads wrote partial lines to HBM).
            const double *__restrict__ dp = a.dp + 3 * q;
            const double *__restrict__ pu = a.pucont + 3 * q;
            double d[3], u[3], o[3];
#pragma unroll
            for (int c = 0; c < 3; c++) { d[c] = dp[c]; u[c] = pu[c]; }
#pragma unroll
            for (int c = 0; c < 3; c++) {
                double rv = zero[c] ? 0.0 : (0.0 + r[c]);
                rv += -1 * d[c];
                o[c] = a.alpha * rv + u[c];
            }
            double *__restrict__ out = a.ucont_out + 3 * q;
#pragma unroll
            for (int c = 0; c < 3; c++) out[c] = o[c];
        } else if (a.cn_r) {
            // Rhs = 0; Rhs += (-1/dt) u; Rhs += (1/dt) u_o; Rhs += 0.5 F (the
            // cell's r); R = Rhs + 0.5 RHS_o - dP (+ F_eul), in the separate
            // passes' operation order (momentum.c:2237-2336)
            const double *__restrict__ u = a.cn_u + 3 * q;
            const double *__restrict__ uo = a.cn_uo + 3 * q;
            const double *__restrict__ ro = a.cn_rhso + 3 * q;
            const double *__restrict__ dp = a.dp + 3 * q;
            double uu[3], uoo[3], roo[3], d[3], fe[3] = {0, 0, 0}, o[3];
#pragma unroll
            for (int c = 0; c < 3; c++) { uu[c] = u[c]; uoo[c] = uo[c]; roo[c] = ro[c]; d[c] = dp[c]; }
            if (a.cn_feul)
#pragma unroll
                for (int c = 0; c < 3; c++) fe[c] = a.cn_feul[3 * q + c];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double s0 = 0.0 + (-a.cn_rdt) * uu[c];
                const double base = s0 + a.cn_rdt * uoo[c];
                const double rv = zero[c] ? 0.0 : base + r[c];
                double s = rv + 0.5 * roo[c];
                s = s + (-1.0) * d[c];
                if (a.cn_feul) s = s + fe[c];
                o[c] = s;
            }
            if (a.cn_rbase)
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const double rb = a.cn_rbase[3 * q + c];
                    o[c] = a.cn_da * rb + a.cn_db * o[c];
                }
            double *__restrict__ out = a.cn_r + 3 * q;
#pragma unroll
            for (int c = 0; c < 3; c++) out[c] = o[c];
        } else {
            for (int c = 0; c < 3; c++) a.rhs[3 * q + c] = zero[c] ? 0.0 : a.rhs[3 * q + c] + r[c];
        }
    }
}

// max |nvert| over every point (ghosts included), one block partial each
__global__ void __launch_bounds__(RED_THREADS) k_nvert_absmax(long long n, const double *__restrict__ nvert,
                                                             double *__restrict__ part)
{
    double v = 0;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        v = fmax(v, fabs(nvert[t]));
    v = block_reduce<OpMax>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

int nvert_all_zero(vfs_ctx *c, bool *zero)
{
    if (c->nvert_zero < 0) {
        hipStream_t st = c->stream;
        hipLaunchKernelGGL(k_nvert_absmax, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, c->g.n, c->f[VFS_NVERT], c->red);
        hipLaunchKernelGGL(finish_reduce<OpMax>, dim3(1), dim3(RED_THREADS), 0, st, c->red, RED_BLOCKS, c->red + RED_BLOCKS);
        VFS_LAUNCH_CHECK("nvert scan");
        VFS_HIP(hipMemcpyAsync(c->red_host, c->red + RED_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, st));
        VFS_HIP(hipStreamSynchronize(st));
        c->nvert_zero = c->red_host[0] == 0.0 ? 1 : 0;
    }
    const char *e = std::getenv("VFS_NOBLANK");
    *zero = c->nvert_zero == 1 && !(e && e[0] == '0');
    return VFS_OK;
}

bool formfunction2_fused(const vfs_ctx *c)
{
    // VFS_FUSED_FP=0 keeps the flux planes (the face pass, then k_cell_fp)
    const char *env = std::getenv("VFS_FUSED_FP");
    return c->p.second_order && c->face_derived && !(env && env[0] == '0');
}

int formfunction2(vfs_ctx *c, double scale, const double *pucont, double rk_alpha, double *ucont_out, const CnResidual *cn)
{
    const Grid &g = c->g;
    hipStream_t st = c->stream;
    const int whole = prof_begin(c, K_FF);
    int rc = periodic_fix(c, c->f[VFS_UCONT], 3, 0);  // momentum.c:638-666
    if (rc) return rc;
    FaceArgs fa;
    fa.g = g;
    fa.ucont = c->f[VFS_UCONT];
    fa.ucat = c->f[VFS_UCAT];
    fa.nut = c->f[VFS_NU_T];
    fa.nvert = c->f[VFS_NVERT];
    fa.fi = {c->f[VFS_ICSI], c->f[VFS_IETA], c->f[VFS_IZET], c->f[VFS_IAJ]};
    fa.fj = {c->f[VFS_JCSI], c->f[VFS_JETA], c->f[VFS_JZET], c->f[VFS_JAJ]};
    fa.fk = {c->f[VFS_KCSI], c->f[VFS_KETA], c->f[VFS_KZET], c->f[VFS_KAJ]};
    fa.cell = {c->f[VFS_CSI], c->f[VFS_ETA], c->f[VFS_ZET], c->f[VFS_AJ]};
    fa.nu = 1. / c->p.ren;
    fa.les = c->p.les;
    fa.second_order = c->p.second_order;
    fa.laplacian = c->p.laplacian;
    fa.out = c->face;
    const unsigned nbt = (unsigned)(((g.mx + FT_I - 1) / FT_I) * ((g.my + FT_J - 1) / FT_J) * ((g.mz + FT_K - 1) / FT_K));
    const long long np[3] = {2LL * g.my * g.mz, 2LL * g.mx * g.mz, 2LL * g.mx * g.my};
    const int per[3] = {g.pi, g.pj, g.pk};
    const bool fused = formfunction2_fused(c);
    if (cn && !fused) return set_error(VFS_EINVAL, "formfunction2: the CN residual needs the fused passes");
    // a grid with no blanked point takes the face pass without blanking tests
    bool nb = false;
    if (fused) {
        int rc = nvert_all_zero(c, &nb);
        if (rc) return rc;
    }
    prof_begin(c, K_FACE);
    if (fused && nb)
        hipLaunchKernelGGL(k_face_fp<true>, dim3(face_fp_blocks(g)), dim3(FT_I * FT_J), 0, st, fa, c->fpc);
    else if (fused)
        hipLaunchKernelGGL(k_face_fp<false>, dim3(face_fp_blocks(g)), dim3(FT_I * FT_J), 0, st, fa, c->fpc);
    else if (fa.second_order && c->face_derived)
        hipLaunchKernelGGL((k_face_fluxes<true, true>), dim3(nbt), dim3(FT_I * FT_J), 0, st, fa);
    else if (fa.second_order)
        hipLaunchKernelGGL((k_face_fluxes<true, false>), dim3(nbt), dim3(FT_I * FT_J), 0, st, fa);
    else
        hipLaunchKernelGGL((k_face_fluxes<false, false>), dim3(nbt), dim3(FT_I * FT_J), 0, st, fa);
    prof_end(c);
    VFS_LAUNCH_CHECK("face fluxes");
    if (fused) {
        // projections of the periodic ghost cells: copies of the cells they stand for
        for (int d = 0; d < 3; d++)
            if (per[d]) hipLaunchKernelGGL(k_periodic_soa, dim3(blocks_for(np[d], FF_TPB)), dim3(FF_TPB), 0, st, g, c->fpc, 3, d);
    } else {
        for (int d = 0; d < 3; d++)
            if (per[d]) hipLaunchKernelGGL(k_periodic_soa, dim3(blocks_for(np[d], FF_TPB)), dim3(FF_TPB), 0, st, g, c->face, 18, d);
    }
    VFS_LAUNCH_CHECK("face periodic");
    CellArgs ca;
    ca.g = g;
    ca.face = c->face;
    ca.csi = c->f[VFS_CSI];
    ca.eta = c->f[VFS_ETA];
    ca.zet = c->f[VFS_ZET];
    ca.iaj = c->f[VFS_IAJ];
    ca.jaj = c->f[VFS_JAJ];
    ca.kaj = c->f[VFS_KAJ];
    ca.nvert = c->f[VFS_NVERT];
    ca.scale = scale;
    ca.second_order = c->p.second_order;
    ca.rhs = c->f[VFS_RHS];
    ca.dp = c->f[VFS_DP];
    ca.pucont = pucont;
    ca.alpha = rk_alpha;
    ca.ucont_out = ucont_out;
    ca.cn_u = c->f[VFS_UCONT];
    ca.cn_uo = cn ? cn->uo : nullptr;
    ca.cn_rhso = cn ? cn->rhso : nullptr;
    ca.cn_feul = cn ? cn->feul : nullptr;
    ca.cn_rdt = cn ? cn->rdt : 0.0;
    ca.cn_r = cn ? cn->r : nullptr;
    ca.cn_rbase = cn ? cn->rbase : nullptr;
    ca.cn_da = cn ? cn->da : 0.0;
    ca.cn_db = cn ? cn->db : 0.0;
    if (!fused) {
        prof_begin(c, K_CELLFP);
        if (ca.second_order)
            hipLaunchKernelGGL(k_cell_fp<true>, dim3(nbt), dim3(FT_I * FT_J), 0, st, ca, c->fpc);
        else
            hipLaunchKernelGGL(k_cell_fp<false>, dim3(nbt), dim3(FT_I * FT_J), 0, st, ca, c->fpc);
        prof_end(c);
    }
    prof_begin(c, K_CELL);
    if (fused && nb) ca.nvert = nullptr;  // the cell pass then skips its blanking tests
    if (fused)
        hipLaunchKernelGGL(k_cell_rhs_pr, dim3(cell_rhs_blocks(g)), dim3(CR_T), 0, st, ca, (const double *)c->fpc);
    else
        hipLaunchKernelGGL(k_cell_rhs, dim3(nbt), dim3(FT_I * FT_J), 0, st, ca, (const double *)c->fpc);
    prof_end(c);
    prof_end(c, whole);
    if (pucont) ucat_stale(c);  // the RK stage wrote Ucont
    VFS_LAUNCH_CHECK("cell rhs");
    return VFS_OK;
}

}  // namespace vfs
