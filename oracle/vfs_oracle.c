/*
 * vfs_oracle.c -- FP64 CPU restatement of the VFS-Wind fractional-step core.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker and the CPU baseline of the
 * bench).  Every function cites the reference lines it restates; the loops
 * follow the reference's order of operations so that results agree to
 * round-off.  The product path (vfs-wind_amd/) never links this file.
 *
 * Scope: the channel / box configurations (DA-periodic directions, no-slip 1,
 * slip 10, outlet 4), -les 0/1, central schemes (second_order 0/1),
 * laplacian 0/1, no level set, no RANS, no immersed boundary, no wall model.
 */
#include "orc_internal.h"
#include <omp.h>

/* OpenMP team size of the oracle's parallel loops (the CPU baseline times it
 * on all cores and on one).  Every reduction keeps a fixed order, so results
 * do not depend on the team size. */
int orc_set_threads(int n)
{
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}

/* The fix reads only cells whose periodic indices are interior, which the fix
 * never writes, so an in-place pass equals the reference's local-array pass. */
void orc_periodic_fix(const orc_params *p, double *f, int nc)
{
    if (!(p->per_i || p->per_j || p->per_k)) return;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < p->mz; k++)
        for (int j = 0; j < p->my; j++)
            for (int i = 0; i < p->mx; i++) {
                int a, b, c, fi, fj, fk;
                if (!periodic_src(p, k, j, i, &c, &b, &a, &fi, &fj, &fk)) continue;
                size_t d = IX(p, k, j, i), s = IX(p, c, b, a);
                for (int q = 0; q < nc; q++) f[nc * d + q] = f[nc * s + q];
            }
}

/* ------------------------------------------------------------------------ */
/* a1: FormMetrics (metrics.c:12-1355)                                       */
/* ------------------------------------------------------------------------ */

#define CX(k, j, i) (coor[3 * IX(p, (k), (j), (i)) + 0])
#define CY(k, j, i) (coor[3 * IX(p, (k), (j), (i)) + 1])
#define CZ(k, j, i) (coor[3 * IX(p, (k), (j), (i)) + 2])

static void cross_metrics(double dxdc, double dydc, double dzdc, double dxde, double dyde,
                          double dzde, double dxdz, double dydz, double dzdz, double *csi,
                          double *eta, double *zet, double *aj)
{
    /* metrics.c:571-589 (and the J/K twins) */
    csi[0] = dyde * dzdz - dzde * dydz;
    csi[1] = -dxde * dzdz + dzde * dxdz;
    csi[2] = dxde * dydz - dyde * dxdz;
    eta[0] = dydz * dzdc - dzdz * dydc;
    eta[1] = -dxdz * dzdc + dzdz * dxdc;
    eta[2] = dxdz * dydc - dydz * dxdc;
    zet[0] = dydc * dzde - dzdc * dyde;
    zet[1] = -dxdc * dzde + dzdc * dxde;
    zet[2] = dxdc * dyde - dydc * dxde;
    double det = dxdc * (dyde * dzdz - dzde * dydz) - dydc * (dxde * dzdz - dzde * dxdz) +
                 dzdc * (dxde * dydz - dyde * dxdz);
    *aj = 1. / det;
}

void orc_form_metrics(const orc_params *p, const double *coor, orc_metrics *m)
{
    const int mx = p->mx, my = p->my, mz = p->mz;
    const size_t n = npts(p);
    memset(m->csi, 0, 3 * n * sizeof(double));
    memset(m->eta, 0, 3 * n * sizeof(double));
    memset(m->zet, 0, 3 * n * sizeof(double));
    memset(m->aj, 0, n * sizeof(double));
    memset(m->cent, 0, 3 * n * sizeof(double));
    double *face_v[9] = {m->icsi, m->ieta, m->izet, m->jcsi, m->jeta, m->jzet,
                         m->kcsi, m->keta, m->kzet};
    double *face_s[3] = {m->iaj, m->jaj, m->kaj};
    for (int q = 0; q < 9; q++) memset(face_v[q], 0, 3 * n * sizeof(double));
    for (int q = 0; q < 3; q++) memset(face_s[q], 0, n * sizeof(double));

    double *csi = m->csi, *eta = m->eta, *zet = m->zet, *aj = m->aj;

    /* cell centres, metrics.c:79-100 */
#pragma omp parallel for schedule(static)
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++)
                for (int c = 0; c < 3; c++) {
                    double s = coor[3 * IX(p, k, j, i) + c] + coor[3 * IX(p, k, j - 1, i) + c] +
                               coor[3 * IX(p, k - 1, j, i) + c] + coor[3 * IX(p, k - 1, j - 1, i) + c] +
                               coor[3 * IX(p, k, j, i - 1) + c] + coor[3 * IX(p, k, j - 1, i - 1) + c] +
                               coor[3 * IX(p, k - 1, j, i - 1) + c] +
                               coor[3 * IX(p, k - 1, j - 1, i - 1) + c];
                    VC(m->cent, c, k, j, i) = 0.125 * s;
                }

    /* The reference compiles FormMetrics with NEWMETRIC defined
     * (metrics.c:10), so the node-plane formulas of metrics.c:115-191 are
     * dead code: the cell metrics csi/eta/zet are the cross products of the
     * cell-averaged edge vectors, computed beside the Jacobian
     * (metrics.c:193-257), and each face metric is the mean of the two cell
     * metrics it separates (metrics.c:588-593, 696-701, 804-809). */
#pragma omp parallel for schedule(static)
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                double d[3][3];
                for (int c = 0; c < 3; c++) {
#define CC(kk, jj, ii) coor[3 * IX(p, (kk), (jj), (ii)) + c]
                    d[c][0] = 0.25 * (CC(k, j, i) + CC(k, j - 1, i) + CC(k - 1, j, i) + CC(k - 1, j - 1, i) -
                                      CC(k, j, i - 1) - CC(k, j - 1, i - 1) - CC(k - 1, j, i - 1) -
                                      CC(k - 1, j - 1, i - 1));
                    d[c][1] = 0.25 * (CC(k, j, i) + CC(k, j, i - 1) + CC(k - 1, j, i) + CC(k - 1, j, i - 1) -
                                      CC(k, j - 1, i) - CC(k, j - 1, i - 1) - CC(k - 1, j - 1, i) -
                                      CC(k - 1, j - 1, i - 1));
                    d[c][2] = 0.25 * (CC(k, j, i) + CC(k, j - 1, i) + CC(k, j, i - 1) + CC(k, j - 1, i - 1) -
                                      CC(k - 1, j, i) - CC(k - 1, j - 1, i) - CC(k - 1, j, i - 1) -
                                      CC(k - 1, j - 1, i - 1));
#undef CC
                }
                size_t q = IX(p, k, j, i);
                cross_metrics(d[0][0], d[1][0], d[2][0], d[0][1], d[1][1], d[2][1], d[0][2], d[1][2], d[2][2],
                              &csi[3 * q], &eta[3 * q], &zet[3 * q], &aj[q]);
            }
    /* ghost-cell mirroring, metrics.c:261-345: each boundary plane copies all
     * four cell arrays from its neighbour plane, in the order i-, i+, j-, j+,
     * k-, k+ (edges and corners take what the later planes copy). */
    double *cell[4] = {csi, eta, zet, aj};
    const int ncomp[4] = {3, 3, 3, 1};
    for (int side = 0; side < 6; side++) {
        const int axis = side >> 1, hi = side & 1;
        const int lim[3] = {mx, my, mz};
        const int to = hi ? lim[axis] - 1 : 0, from = hi ? lim[axis] - 2 : 1;
        for (int k = 0; k < mz; k++)
            for (int j = 0; j < my; j++)
                for (int i = 0; i < mx; i++) {
                    int ijk[3] = {i, j, k};
                    if (ijk[axis] != to) continue;
                    ijk[axis] = from;
                    size_t d = IX(p, k, j, i), s = IX(p, ijk[2], ijk[1], ijk[0]);
                    for (int a = 0; a < 4; a++)
                        for (int c = 0; c < ncomp[a]; c++) cell[a][ncomp[a] * d + c] = cell[a][ncomp[a] * s + c];
                }
    }
    /* periodic copy into the ghosted cell arrays before the face averages
     * read them, metrics.c:365-395 */
    orc_periodic_fix(p, csi, 3);
    orc_periodic_fix(p, eta, 3);
    orc_periodic_fix(p, zet, 3);
    orc_periodic_fix(p, aj, 1);

    /* face metrics: F = 0.5 X + 0.5 Y (AxByC, momentum.c:40-44) of the two
     * cells, and the harmonic mean of their aj; I faces i = 0..mx-2 over the
     * interior j, k (metrics.c:497, 588-593), J faces (626, 696-701), K faces
     * (731, 804-809). */
    for (int axis = 0; axis < 3; axis++) {
        double *fv[3] = {face_v[3 * axis], face_v[3 * axis + 1], face_v[3 * axis + 2]};
        double *fa = face_s[axis];
        const size_t step = axis == 0 ? 1 : axis == 1 ? (size_t)mx : (size_t)mx * my;
#pragma omp parallel for schedule(static)
        for (int k = (axis == 2 ? 0 : 1); k < mz - 1; k++)
            for (int j = (axis == 1 ? 0 : 1); j < my - 1; j++)
                for (int i = (axis == 0 ? 0 : 1); i < mx - 1; i++) {
                    size_t q = IX(p, k, j, i), r = q + step;
                    for (int a = 0; a < 3; a++)
                        for (int c = 0; c < 3; c++)
                            fv[a][3 * q + c] = 0.5 * cell[a][3 * q + c] + 0.5 * cell[a][3 * r + c];
                    fa[q] = 2. / (1. / aj[q] + 1. / aj[r]);
                }
    }
    /* final periodic fix of all sixteen arrays, metrics.c:1180-1300 */
    orc_periodic_fix(p, m->csi, 3);
    orc_periodic_fix(p, m->eta, 3);
    orc_periodic_fix(p, m->zet, 3);
    orc_periodic_fix(p, m->aj, 1);
    for (int q = 0; q < 9; q++) orc_periodic_fix(p, face_v[q], 3);
    for (int q = 0; q < 3; q++) orc_periodic_fix(p, face_s[q], 1);
}
#undef CX
#undef CY
#undef CZ

/* ------------------------------------------------------------------------ */
/* a2: Contra2Cart_2 (rhs.c:65-749)                                          */
/* ------------------------------------------------------------------------ */

/* Cramer's rule solve, rhs.c:215-239 */
static void cramer(const double *csi, const double *eta, const double *zet, const double q[3],
                   double *u)
{
    double m00 = csi[0], m01 = csi[1], m02 = csi[2];
    double m10 = eta[0], m11 = eta[1], m12 = eta[2];
    double m20 = zet[0], m21 = zet[1], m22 = zet[2];
    double det = m00 * (m11 * m22 - m12 * m21) - m01 * (m10 * m22 - m12 * m20) + m02 * (m10 * m21 - m11 * m20);
    double det0 = q[0] * (m11 * m22 - m12 * m21) - q[1] * (m01 * m22 - m02 * m21) + q[2] * (m01 * m12 - m02 * m11);
    double det1 = -q[0] * (m10 * m22 - m12 * m20) + q[1] * (m00 * m22 - m02 * m20) - q[2] * (m00 * m12 - m02 * m10);
    double det2 = q[0] * (m10 * m21 - m11 * m20) - q[1] * (m00 * m21 - m01 * m20) + q[2] * (m00 * m11 - m01 * m10);
    u[0] = det0 / det;
    u[1] = det1 / det;
    u[2] = det2 / det;
}

/* Inverse of the contravariant matrix g = [csi; eta; zet]: G[r][c] = dx_r/dxi_c. */
void orc_covariant(const double g[3][3], double G[3][3])
{
    /* Calculate_Covariant_metrics, rhs2.c:595-612 */
    const double a11 = g[0][0], a12 = g[0][1], a13 = g[0][2];
    const double a21 = g[1][0], a22 = g[1][1], a23 = g[1][2];
    const double a31 = g[2][0], a32 = g[2][1], a33 = g[2][2];
    double det = a11 * (a33 * a22 - a32 * a23) - a21 * (a33 * a12 - a32 * a13) + a31 * (a23 * a12 - a22 * a13);
    G[0][0] = (a33 * a22 - a32 * a23) / det, G[0][1] = -(a33 * a12 - a32 * a13) / det, G[0][2] = (a23 * a12 - a22 * a13) / det;
    G[1][0] = -(a33 * a21 - a31 * a23) / det, G[1][1] = (a33 * a11 - a31 * a13) / det, G[1][2] = -(a23 * a11 - a21 * a13) / det;
    G[2][0] = (a32 * a21 - a31 * a22) / det, G[2][1] = -(a32 * a11 - a31 * a12) / det, G[2][2] = (a22 * a11 - a21 * a12) / det;
}

/* Slip reflection about the face normal taken from the covariant metric of
 * the first interior cell, rhs.c:447-544.  dir: 0/1/2 = i/j/k column of G,
 * sgn: -1 on the low face, +1 on the high face. */
static void slip_fill(const orc_params *p, const orc_metrics *m, const double *ucat_old, int kn,
                      int jn, int in, int dir, double sgn, double *out)
{
    double g[3][3], G[3][3];
    for (int c = 0; c < 3; c++) {
        g[0][c] = VC(m->csi, c, kn, jn, in);
        g[1][c] = VC(m->eta, c, kn, jn, in);
        g[2][c] = VC(m->zet, c, kn, jn, in);
    }
    orc_covariant(g, G);
    double nx = sgn * G[0][dir], ny = sgn * G[1][dir], nz = sgn * G[2][dir];
    double sum = sqrt(nx * nx + ny * ny + nz * nz);
    nx /= sum, ny /= sum, nz /= sum;
    double ux = VX(ucat_old, kn, jn, in), uy = VY(ucat_old, kn, jn, in), uz = VZ(ucat_old, kn, jn, in);
    double un = ux * nx + uy * ny + uz * nz;
    out[0] = ux - 2 * un * nx;
    out[1] = uy - 2 * un * ny;
    out[2] = uz - 2 * un * nz;
}

void orc_contra2cart(const orc_params *p, const orc_metrics *m, double *ucont, double *ucat,
                     const double *ucat_old)
{
    const int mx = p->mx, my = p->my, mz = p->mz;
    const double *nvert = m->nvert;
    const size_t n = npts(p);

    /* rhs.c:129-156 */
    orc_periodic_fix(p, ucont, 3);
    /* rhs.c:158-247 */
#pragma omp parallel for schedule(static)
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                if (S(nvert, k, j, i) >= 0.1) continue;
                double q[3];
                q[0] = 0.5 * (VX(ucont, k, j, i - 1) + VX(ucont, k, j, i));
                q[1] = 0.5 * (VY(ucont, k, j - 1, i) + VY(ucont, k, j, i));
                q[2] = 0.5 * (VZ(ucont, k - 1, j, i) + VZ(ucont, k, j, i));
                size_t c = IX(p, k, j, i);
                cramer(&m->csi[3 * c], &m->eta[3 * c], &m->zet[3 * c], q, &ucat[3 * c]);
            }
    /* rhs.c:254-291: periodic fix of ucat */
    orc_periodic_fix(p, ucat, 3);
    /* rhs.c:300-680: ghost-cell BC fill; reads the snapshot L2 */
    double *L2 = dalloc(3 * n);
    memcpy(L2, ucat, 3 * n * sizeof(double));
    const int *bc = p->bctype;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                double *u = &ucat[3 * IX(p, k, j, i)];
                int solid_flag = 0;
                if ((int)(S(nvert, k, j, i) + 0.1) == 3) {
                    u[0] = u[1] = u[2] = 0;
                    continue;
                }
                /* slip, rhs.c:447-544 */
                if (bc[0] == 10 && i == 0 && (j != 0 && k != 0)) {
                    slip_fill(p, m, ucat_old, k, j, i + 1, 0, -1.0, u);
                    if (S(nvert, k, j, i + 1) > 0.1) u[0] = u[1] = u[2] = 0;
                    if (solid_flag) u[0] = u[1] = u[2] = 0;
                }
                if (bc[1] == 10 && i == mx - 1 && (j != 0 && k != 0)) {
                    slip_fill(p, m, ucat_old, k, j, i - 1, 0, 1.0, u);
                    if (S(nvert, k, j, i - 1) > 0.1) u[0] = u[1] = u[2] = 0;
                    if (solid_flag) u[0] = u[1] = u[2] = 0;
                }
                if (bc[2] == 10 && j == 0 && (i != 0 && k != 0)) {
                    slip_fill(p, m, ucat_old, k, j + 1, i, 1, -1.0, u);
                    if (S(nvert, k, j + 1, i) > 0.1) u[0] = u[1] = u[2] = 0;
                    if (solid_flag) u[0] = u[1] = u[2] = 0;
                }
                if (abs(bc[3]) == 10 && j == my - 1 && (i != 0 && k != 0)) {
                    slip_fill(p, m, ucat_old, k, j - 1, i, 1, 1.0, u);
                    if (S(nvert, k, j - 1, i) > 0.1) u[0] = u[1] = u[2] = 0;
                    if (solid_flag) u[0] = u[1] = u[2] = 0;
                }
                /* no-slip, rhs.c:547-590 */
                if (i == 0 && bc[0] == 1 && (j != 0 && k != 0)) {
                    for (int c = 0; c < 3; c++) u[c] = -VC(L2, c, k, j, i + 1);
                    solid_flag = 1;
                }
                if (i == mx - 1 && bc[1] == 1 && (j != 0 && k != 0)) {
                    for (int c = 0; c < 3; c++) u[c] = -VC(L2, c, k, j, i - 1);
                    solid_flag = 1;
                }
                if (j == 0 && bc[2] == 1 && (i != 0 && k != 0)) {
                    for (int c = 0; c < 3; c++) u[c] = -VC(L2, c, k, j + 1, i);
                    solid_flag = 1;
                }
                if (j == my - 1 && bc[3] == 1 && (i != 0 && k != 0)) {
                    for (int c = 0; c < 3; c++) u[c] = -VC(L2, c, k, j - 1, i);
                    solid_flag = 1;
                }
                if (k == 0 && bc[4] == 1 && (i != 0 && j != 0)) {
                    for (int c = 0; c < 3; c++) u[c] = -VC(L2, c, k + 1, j, i);
                    solid_flag = 1;
                }
                if (k == mz - 1 && bc[5] == 1 && (i != 0 && j != 0)) {
                    for (int c = 0; c < 3; c++) u[c] = -VC(L2, c, k - 1, j, i);
                    solid_flag = 1;
                }
                /* lid-driven cavity j+, rhs.c:593-605: the wall moves with u = 1 */
                if (j == my - 1 && bc[3] == 2) {
                    const double *nb = solid_flag ? &VX(ucat_old, k, j - 1, i) : &VX(L2, k, j - 1, i);
                    u[0] = 2.0 - nb[0];
                    u[1] = -nb[1];
                    u[2] = -nb[2];
                }
                /* Couette walls j- and j+, rhs.c:608-624: the wall moves with w = 1 */
                if (j == 0 && bc[2] == 12) {
                    u[0] = -VX(L2, k, j + 1, i);
                    u[1] = -VY(L2, k, j + 1, i);
                    u[2] = 2.0 - VZ(L2, k, j + 1, i);
                    if (solid_flag) u[0] = u[1] = u[2] = 0;
                }
                if (j == my - 1 && bc[3] == 12) {
                    u[0] = -VX(L2, k, j - 1, i);
                    u[1] = -VY(L2, k, j - 1, i);
                    u[2] = 2.0 - VZ(L2, k, j - 1, i);
                    if (solid_flag) u[0] = u[1] = u[2] = 0;
                }
                /* outflow j+, rhs.c:640-642 */
                if (bc[3] == 4 && j == my - 1 && (i != 0 && i != mx - 1 && k != 0 && k != mz - 1))
                    for (int c = 0; c < 3; c++) u[c] = VC(L2, c, k, j - 1, i);
                /* outlet k+, rhs.c:641-660 */
                if (bc[5] == 4 && k == mz - 1 && (i != 0 && i != mx - 1 && j != 0 && j != my - 1)) {
                    if (S(nvert, k - 1, j, i) > 0.1) u[0] = u[1] = u[2] = 0;
                    if (solid_flag) u[0] = u[1] = u[2] = 0;
                }
                /* corner cells next to two walls, rhs.c:664-669 */
                if (i != 0 && i != mx - 1 && j != 0 && j != my - 2 && k != 0 && k != mz - 2) {
                    if (bc[0] <= 1 && bc[2] <= 1 && (i == 1 && j == 1)) u[0] = u[1] = u[2] = 0;
                    if (bc[1] <= 1 && bc[2] <= 1 && (i == mx - 2 && j == 1)) u[0] = u[1] = u[2] = 0;
                }
            }
    /* rhs.c:712-748: periodic fix from the post-BC snapshot L3 */
    orc_periodic_fix(p, ucat, 3);
    free(L2);
}

/* ------------------------------------------------------------------------ */
/* FormBCS subset (bcs.c:1684-2767)                                          */
/* ------------------------------------------------------------------------ */

void orc_form_bcs(const orc_params *p, const orc_metrics *m, double *ucont, double *ucat,
                  const double *ucat_old, int outflow_scale)
{
    const int mx = p->mx, my = p->my, mz = p->mz;
    const double *nvert = m->nvert;
    const double threshold = 0.1; /* bcs.c:11 */
    orc_contra2cart(p, m, ucont, ucat, ucat_old);
    const int *bc = p->bctype;

    /* far field (code 6), bcs.c:1745-1993, no frame motion: the boundary
     * fluxes from the adjacent cell's velocity; with a far-field k+ face the
     * in/out flux difference is spread over the far-field faces */
    double FarFluxIn = 0, FarFluxOut = 0, FarAreaIn = 0, FarAreaOut = 0;
    if (bc[0] == 6)
        for (int k = 1; k < mz - 1; k++)
            for (int j = 1; j < my - 1; j++) {
                VX(ucont, k, j, 0) = VX(ucat, k, j, 1) * VX(m->csi, k, j, 1);
                FarFluxIn += VX(ucont, k, j, 0);
                FarAreaIn += VX(m->csi, k, j, 0);
            }
    if (bc[1] == 6)
        for (int k = 1; k < mz - 1; k++)
            for (int j = 1; j < my - 1; j++) {
                VX(ucont, k, j, mx - 2) = VX(ucat, k, j, mx - 2) * VX(m->csi, k, j, mx - 2);
                FarFluxOut += VX(ucont, k, j, mx - 2);
                FarAreaOut += VX(m->csi, k, j, mx - 2);
            }
    if (bc[2] == 6)
        for (int k = 1; k < mz - 1; k++)
            for (int i = 1; i < mx - 1; i++) {
                VY(ucont, k, 0, i) = VY(ucat, k, 1, i) * VY(m->eta, k, 1, i);
                FarFluxIn += VY(ucont, k, 0, i);
                FarAreaIn += VY(m->eta, k, 0, i);
            }
    if (bc[3] == 6)
        for (int k = 1; k < mz - 1; k++)
            for (int i = 1; i < mx - 1; i++) {
                VY(ucont, k, my - 2, i) = VY(ucat, k, my - 2, i) * VY(m->eta, k, my - 2, i);
                FarFluxOut += VY(ucont, k, my - 2, i);
                FarAreaOut += VY(m->eta, k, my - 2, i);
            }
    if (bc[4] == 6)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                VZ(ucont, 0, j, i) = VZ(ucat, 1, j, i) * VZ(m->zet, 1, j, i);
                FarFluxIn += VZ(ucont, 0, j, i);
                FarAreaIn += VZ(m->zet, 0, j, i);
            }
    if (bc[5] == 6)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                VZ(ucont, mz - 2, j, i) = VZ(ucat, mz - 2, j, i) * VZ(m->zet, mz - 2, j, i);
                FarFluxOut += VZ(ucont, mz - 2, j, i);
                FarAreaOut += VZ(m->zet, mz - 2, j, i);
            }
    if (bc[5] == 6) {
        const double FluxDiff = 0.5 * (FarFluxIn - FarFluxOut);
        double VelDiffIn = FluxDiff / FarAreaIn, VelDiffOut = FluxDiff / FarAreaOut;
        if (fabs(FluxDiff) < 1.e-6 || fabs(FarAreaIn) < 1.e-6) VelDiffIn = 0.;
        if (fabs(FluxDiff) < 1.e-6 || fabs(FarAreaOut) < 1.e-6) VelDiffOut = 0.;
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++)
                VZ(ucont, mz - 2, j, i) = (VZ(ucat, mz - 2, j, i) + VelDiffOut) * VZ(m->zet, mz - 2, j, i);
        if (bc[3] == 6)
            for (int k = 1; k < mz - 1; k++)
                for (int i = 1; i < mx - 1; i++)
                    VY(ucont, k, my - 2, i) = (VY(ucat, k, my - 2, i) + VelDiffOut) * VY(m->eta, k, my - 2, i);
        if (bc[1] == 6)
            for (int k = 1; k < mz - 1; k++)
                for (int j = 1; j < my - 1; j++)
                    VX(ucont, k, j, mx - 2) = (VX(ucat, k, j, mx - 2) + VelDiffOut) * VX(m->csi, k, j, mx - 2);
        if (bc[0] == 6)
            for (int k = 1; k < mz - 1; k++)
                for (int j = 1; j < my - 1; j++) VX(ucont, k, j, 0) = (VX(ucat, k, j, 1) - VelDiffIn) * VX(m->csi, k, j, 1);
        if (bc[2] == 6)
            for (int k = 1; k < mz - 1; k++)
                for (int i = 1; i < mx - 1; i++) VY(ucont, k, 0, i) = (VY(ucat, k, 1, i) - VelDiffIn) * VY(m->eta, k, 1, i);
        if (bc[4] == 6)
            for (int j = 1; j < my - 1; j++)
                for (int i = 1; i < mx - 1; i++) VZ(ucont, 0, j, i) = (VZ(ucat, 1, j, i) - VelDiffIn) * VZ(m->zet, 1, j, i);
    }

    /* characteristic outlet (code 8) on k+, bcs.c:1997-2041: after the first
     * two steps the outlet fluxes are scaled to carry the inflow */
    if (bc[5] == 8) {
        double FluxOut = 0;
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) FluxOut += VZ(ucont, mz - 2, j, i);
        const double FluxIn = p->flux_in_sum + FarFluxIn;
        double ratio = FluxIn / FluxOut;
        if (fabs(FluxOut) < 1.e-6) ratio = 1.;
        if (fabs(FluxIn) < 1.e-6) ratio = 0.;
        if (!(p->ti == 0 || p->ti == 1))
            for (int j = 1; j < my - 1; j++)
                for (int i = 1; i < mx - 1; i++) VZ(ucont, mz - 2, j, i) = VZ(ucont, mz - 2, j, i) * ratio;
    }

    /* outlet flux balance, bcs.c:2127-2260 */
    if (p->bctype[5] == 4) {
        int k = mz - 1;
        double FluxOut = 0, lArea = 0;
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++)
                if (S(nvert, k - 1, j, i) < threshold) {
                    FluxOut += VZ(ucont, k - 1, j, i);
                    double zx = VX(m->zet, k - 1, j, i), zy = VY(m->zet, k - 1, j, i), zz = VZ(m->zet, k - 1, j, i);
                    lArea += sqrt(zx * zx + zy * zy + zz * zz);
                }
        if (outflow_scale) {
            for (int j = 1; j < my - 1; j++)
                for (int i = 1; i < mx - 1; i++) {
                    double zx = VX(m->zet, k - 1, j, i), zy = VY(m->zet, k - 1, j, i), zz = VZ(m->zet, k - 1, j, i);
                    double Area = sqrt(zx * zx + zy * zy + zz * zz);
                    if (S(nvert, k - 1, j, i) < threshold)
                        VZ(ucont, k - 1, j, i) += (p->flux_in_sum - FluxOut) * Area / lArea;
                }
        }
    }
    /* outlet (code 4) on k-, bcs.c:2294-2340 */
    if (bc[4] == 4) {
        double FluxOut = 0, lArea = 0;
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++)
                if (S(nvert, 0, j, i) < threshold) {
                    FluxOut += (VX(ucat, 1, j, i) * (VX(m->zet, 0, j, i) + VX(m->zet, 1, j, i)) +
                                VY(ucat, 1, j, i) * (VY(m->zet, 0, j, i) + VY(m->zet, 1, j, i)) +
                                VZ(ucat, 1, j, i) * (VZ(m->zet, 0, j, i) + VZ(m->zet, 1, j, i))) * 0.5;
                    lArea += VZ(m->zet, 0, j, i);
                }
        const double ratio = (p->flux_in_sum - FluxOut) / lArea;
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) VZ(ucont, 0, j, i) = (VZ(ucat, 1, j, i) + ratio) * VZ(m->zet, 0, j, i);
    }
    /* wall normal flux, bcs.c:2347-2386 */
    if (p->bctype[0] == 1 || p->bctype[0] == 10)
        for (int k = 1; k < mz - 1; k++)
            for (int j = 1; j < my - 1; j++) VX(ucont, k, j, 0) = 0;
    if (p->bctype[1] == 1 || p->bctype[1] == 10)
        for (int k = 1; k < mz - 1; k++)
            for (int j = 1; j < my - 1; j++) VX(ucont, k, j, mx - 2) = 0;
    if (p->bctype[2] == 1 || p->bctype[2] == 10)
        for (int k = 1; k < mz - 1; k++)
            for (int i = 1; i < mx - 1; i++) VY(ucont, k, 0, i) = 0;
    if (p->bctype[3] == 1 || p->bctype[3] == 10)
        for (int k = 1; k < mz - 1; k++)
            for (int i = 1; i < mx - 1; i++) VY(ucont, k, my - 2, i) = 0;

    orc_contra2cart(p, m, ucont, ucat, ucat_old);

    /* zero the twelve edges of the ghost layer, bcs.c:2630-2760 */
    for (int j = 0; j < my; j++) {
        for (int c = 0; c < 3; c++) {
            VC(ucat, c, 0, j, 0) = 0; VC(ucat, c, 0, j, mx - 1) = 0;
            VC(ucat, c, mz - 1, j, 0) = 0; VC(ucat, c, mz - 1, j, mx - 1) = 0;
        }
    }
    for (int i = 0; i < mx; i++) {
        for (int c = 0; c < 3; c++) {
            VC(ucat, c, 0, 0, i) = 0; VC(ucat, c, 0, my - 1, i) = 0;
            VC(ucat, c, mz - 1, 0, i) = 0; VC(ucat, c, mz - 1, my - 1, i) = 0;
        }
    }
    for (int k = 0; k < mz; k++) {
        for (int c = 0; c < 3; c++) {
            VC(ucat, c, k, 0, 0) = 0; VC(ucat, c, k, 0, mx - 1) = 0;
            VC(ucat, c, k, my - 1, 0) = 0; VC(ucat, c, k, my - 1, mx - 1) = 0;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* a4: Compute_eddy_viscosity_LES (les.c:1143-1361), Compute_du_center       */
/* (k-omega.c:313-431), Compute_du_dxyz (k-omega.c:605-618)                  */
/* ------------------------------------------------------------------------ */

static const double SOLID = 0.1; /* k-omega.c:19 */

void orc_du_center(const orc_params *p, const double *nvert, const double *ucat, int k, int j, int i,
                   double d[3][3])
{
    /* d[c][0] = d u_c / d csi, d[c][1] = d/d eta, d[c][2] = d/d zeta */
    const int mx = p->mx, my = p->my, mz = p->mz;
    for (int c = 0; c < 3; c++) {
        if (S(nvert, k, j, i + 1) > SOLID || (!p->per_i && i == mx - 2))
            d[c][0] = (VC(ucat, c, k, j, i) - VC(ucat, c, k, j, i - 1));
        else if (S(nvert, k, j, i - 1) > SOLID || (!p->per_i && i == 1))
            d[c][0] = (VC(ucat, c, k, j, i + 1) - VC(ucat, c, k, j, i));
        else
            d[c][0] = (VC(ucat, c, k, j, i + 1) - VC(ucat, c, k, j, i - 1)) * 0.5;
        if (S(nvert, k, j + 1, i) > SOLID || (!p->per_j && j == my - 2))
            d[c][1] = (VC(ucat, c, k, j, i) - VC(ucat, c, k, j - 1, i));
        else if (S(nvert, k, j - 1, i) > SOLID || (!p->per_j && j == 1))
            d[c][1] = (VC(ucat, c, k, j + 1, i) - VC(ucat, c, k, j, i));
        else
            d[c][1] = (VC(ucat, c, k, j + 1, i) - VC(ucat, c, k, j - 1, i)) * 0.5;
        /* k-omega.c:390-395 tests k==0 (not k==1) for the lower one-sided case */
        if (S(nvert, k + 1, j, i) > SOLID || (!p->per_k && k == mz - 2))
            d[c][2] = (VC(ucat, c, k, j, i) - VC(ucat, c, k - 1, j, i));
        else if (S(nvert, k - 1, j, i) > SOLID || (!p->per_k && k == 0))
            d[c][2] = (VC(ucat, c, k + 1, j, i) - VC(ucat, c, k, j, i));
        else
            d[c][2] = (VC(ucat, c, k + 1, j, i) - VC(ucat, c, k - 1, j, i)) * 0.5;
    }
}

void orc_eddy_viscosity(const orc_params *p, const orc_metrics *m, const double *ucat, double *nu_t)
{
    orc_eddy_viscosity_cs(p, m, ucat, NULL, nu_t);
}

void orc_eddy_viscosity_cs(const orc_params *p, const orc_metrics *m, const double *ucat, const double *lcs,
                           double *nu_t)
{
    const int mx = p->mx, my = p->my, mz = p->mz;
    memset(nu_t, 0, npts(p) * sizeof(double));
    if (!p->les) return;
#pragma omp parallel for schedule(static)
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                if (S(m->nvert, k, j, i) > 1.1) { S(nu_t, k, j, i) = 0; continue; }
                size_t q = IX(p, k, j, i);
                double ajc = m->aj[q];
                const double *cs = &m->csi[3 * q], *et = &m->eta[3 * q], *ze = &m->zet[3 * q];
                double d[3][3];
                orc_du_center(p, m->nvert, ucat, k, j, i, d);
                /* Compute_du_dxyz: du_c/dx_r = (d[c][0]*csi_r + d[c][1]*eta_r + d[c][2]*zet_r)*aj */
                double g[3][3];
                for (int c = 0; c < 3; c++)
                    for (int r = 0; r < 3; r++)
                        g[c][r] = (d[c][0] * cs[r] + d[c][1] * et[r] + d[c][2] * ze[r]) * ajc;
                double Sxx = 0.5 * (g[0][0] + g[0][0]), Sxy = 0.5 * (g[0][1] + g[1][0]), Sxz = 0.5 * (g[0][2] + g[2][0]);
                double Syx = Sxy, Syy = 0.5 * (g[1][1] + g[1][1]), Syz = 0.5 * (g[1][2] + g[2][1]);
                double Szx = Sxz, Szy = Syz, Szz = 0.5 * (g[2][2] + g[2][2]);
                double Sabs = sqrt(2.0 * (Sxx * Sxx + Sxy * Sxy + Sxz * Sxz + Syx * Syx + Syy * Syy + Syz * Syz +
                                          Szx * Szx + Szy * Szy + Szz * Szz));
                double filter = pow(1. / ajc, 1. / 3.);
                S(nu_t, k, j, i) = (lcs ? lcs[q] : p->cs2) * pow(filter, 2.0) * Sabs;
            }
    orc_periodic_fix(p, nu_t, 1);
}

/* ------------------------------------------------------------------------ */
/* Calc_k_Flux (rhs2.c:1549-1633) and mean_k_area (bcs.c:137-199)            */
/* ------------------------------------------------------------------------ */

double orc_calc_k_flux(const orc_params *p, const orc_metrics *m, const double *ucont)
{
    double flux = 0;
    for (int k = 1; k < p->mz - 1; k++)
        for (int j = 1; j < p->my - 1; j++)
            for (int i = 1; i < p->mx - 1; i++)
                if (S(m->nvert, k - 1, j, i) + S(m->nvert, k, j, i) < 0.1) flux += VZ(ucont, k, j, i);
    return flux / (p->mz - 2);
}

double orc_mean_k_area(const orc_params *p, const orc_metrics *m)
{
    double area = 0;
    for (int k = 1; k < p->mz - 1; k++) {
        double a = 0;
        for (int j = 1; j < p->my - 1; j++)
            for (int i = 1; i < p->mx - 1; i++)
                if (S(m->nvert, k - 1, j, i) + S(m->nvert, k, j, i) < 0.1) {
                    double zx = VX(m->zet, k, j, i), zy = VY(m->zet, k, j, i), zz = VZ(m->zet, k, j, i);
                    a += sqrt(zx * zx + zy * zy + zz * zz);
                }
        area += a;
    }
    return area / (p->mz - 2);
}

/* ------------------------------------------------------------------------ */
/* a5: Pressure_Gradient (momentum.c:203-439)                                */
/* ------------------------------------------------------------------------ */

static inline int nv1(const orc_params *p, const double *nvert, int k, int j, int i)
{
    return (int)(S(nvert, k, j, i) + 0.5) == 1;
}

void orc_pressure_gradient(const orc_params *p, const orc_metrics *m, double *pres, double *dp)
{
    const int mx = p->mx, my = p->my, mz = p->mz;
    const double *nv = m->nvert;
    orc_periodic_fix(p, pres, 1); /* momentum.c:249-279 */
    memset(dp, 0, 3 * npts(p) * sizeof(double));
    const int pi_ = p->per_i, pj = p->per_j, pk = p->per_k;
#pragma omp parallel for schedule(static)
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t q = IX(p, k, j, i);
                const double *ic = &m->icsi[3 * q], *ie = &m->ieta[3 * q], *iz = &m->izet[3 * q];
                const double *jc = &m->jcsi[3 * q], *je = &m->jeta[3 * q], *jz = &m->jzet[3 * q];
                const double *kc = &m->kcsi[3 * q], *ke = &m->keta[3 * q], *kz = &m->kzet[3 * q];
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
#define P(kk, jj, ii) S(pres, kk, jj, ii)
                /* i face, momentum.c:345-361 */
                if (i == mx - 2 && pi_) dpdc = P(k, j, mx + 1) - P(k, j, i);
                else dpdc = P(k, j, i + 1) - P(k, j, i);
                if (nv1(p, nv, k, j + 1, i) || nv1(p, nv, k, j + 1, i + 1) || (j == my - 2 && !pj))
                    dpde = (P(k, j, i) - P(k, j - 1, i) + P(k, j, i + 1) - P(k, j - 1, i + 1)) * 0.5;
                else if (nv1(p, nv, k, j - 1, i) || nv1(p, nv, k, j - 1, i + 1) || (j == 1 && !pj))
                    dpde = (P(k, j + 1, i) - P(k, j, i) + P(k, j + 1, i + 1) - P(k, j, i + 1)) * 0.5;
                else
                    dpde = (P(k, j + 1, i) - P(k, j - 1, i) + P(k, j + 1, i + 1) - P(k, j - 1, i + 1)) * 0.25;
                if (nv1(p, nv, k + 1, j, i) || nv1(p, nv, k + 1, j, i + 1) || (k == mz - 2 && !pk))
                    dpdz = (P(k, j, i) - P(k - 1, j, i) + P(k, j, i + 1) - P(k - 1, j, i + 1)) * 0.5;
                else if (nv1(p, nv, k - 1, j, i) || nv1(p, nv, k - 1, j, i + 1) || (k == 1 && !pk))
                    dpdz = (P(k + 1, j, i) - P(k, j, i) + P(k + 1, j, i + 1) - P(k, j, i + 1)) * 0.5;
                else
                    dpdz = (P(k + 1, j, i) - P(k - 1, j, i) + P(k + 1, j, i + 1) - P(k - 1, j, i + 1)) * 0.25;
                VX(dp, k, j, i) = (dpdc * g11_i + dpde * g12_i + dpdz * g13_i) * m->iaj[q];
                /* j face, momentum.c:363-375 */
                if (nv1(p, nv, k, j, i + 1) || nv1(p, nv, k, j + 1, i + 1) || (i == mx - 2 && !pi_))
                    dpdc = (P(k, j, i) - P(k, j, i - 1) + P(k, j + 1, i) - P(k, j + 1, i - 1)) * 0.5;
                else if (nv1(p, nv, k, j, i - 1) || nv1(p, nv, k, j + 1, i - 1) || (i == 1 && !pi_))
                    dpdc = (P(k, j, i + 1) - P(k, j, i) + P(k, j + 1, i + 1) - P(k, j + 1, i)) * 0.5;
                else
                    dpdc = (P(k, j, i + 1) - P(k, j, i - 1) + P(k, j + 1, i + 1) - P(k, j + 1, i - 1)) * 0.25;
                if (j == my - 2 && pj) dpde = P(k, my + 1, i) - P(k, j, i);
                else dpde = P(k, j + 1, i) - P(k, j, i);
                if (nv1(p, nv, k + 1, j, i) || nv1(p, nv, k + 1, j + 1, i) || (k == mz - 2 && !pk))
                    dpdz = (P(k, j, i) - P(k - 1, j, i) + P(k, j + 1, i) - P(k - 1, j + 1, i)) * 0.5;
                else if (nv1(p, nv, k - 1, j, i) || nv1(p, nv, k - 1, j + 1, i) || (k == 1 && !pk))
                    dpdz = (P(k + 1, j, i) - P(k, j, i) + P(k + 1, j + 1, i) - P(k, j + 1, i)) * 0.5;
                else
                    dpdz = (P(k + 1, j, i) - P(k - 1, j, i) + P(k + 1, j + 1, i) - P(k - 1, j + 1, i)) * 0.25;
                VY(dp, k, j, i) = (dpdc * g21_j + dpde * g22_j + dpdz * g23_j) * m->jaj[q];
                /* k face, momentum.c:377-401 */
                if (nv1(p, nv, k, j, i + 1) || nv1(p, nv, k + 1, j, i + 1) || (i == mx - 2 && !pi_))
                    dpdc = (P(k, j, i) - P(k, j, i - 1) + P(k + 1, j, i) - P(k + 1, j, i - 1)) * 0.5;
                else if (nv1(p, nv, k, j, i - 1) || nv1(p, nv, k + 1, j, i - 1) || (i == 1 && !pi_))
                    dpdc = (P(k, j, i + 1) - P(k, j, i) + P(k + 1, j, i + 1) - P(k + 1, j, i)) * 0.5;
                else
                    dpdc = (P(k, j, i + 1) - P(k, j, i - 1) + P(k + 1, j, i + 1) - P(k + 1, j, i - 1)) * 0.25;
                if (nv1(p, nv, k, j + 1, i) || nv1(p, nv, k + 1, j + 1, i) || (j == my - 2 && !pj))
                    dpde = (P(k, j, i) - P(k, j - 1, i) + P(k + 1, j, i) - P(k + 1, j - 1, i)) * 0.5;
                else if (nv1(p, nv, k, j - 1, i) || nv1(p, nv, k + 1, j - 1, i) || (j == 1 && !pj))
                    dpde = (P(k, j + 1, i) - P(k, j, i) + P(k + 1, j + 1, i) - P(k + 1, j, i)) * 0.5;
                else
                    dpde = (P(k, j + 1, i) - P(k, j - 1, i) + P(k + 1, j + 1, i) - P(k + 1, j - 1, i)) * 0.25;
                if (k == mz - 2 && pk) dpdz = P(mz + 1, j, i) - P(k, j, i);
                else dpdz = (P(k + 1, j, i) - P(k, j, i));
                if (pk) {
                    double dz = 1. / m->kaj[q] / sqrt(kz[0] * kz[0] + kz[1] * kz[1] + kz[2] * kz[2]);
                    dpdz += dz * (p->mean_k_flux - p->inlet_flux) / p->dt / p->mean_k_area;
                }
                VZ(dp, k, j, i) = (dpdc * g31_k + dpde * g32_k + dpdz * g33_k) * m->kaj[q];
#undef P
                /* momentum.c:403-411 */
                if (i == 0 || S(nv, k, j, i) + S(nv, k, j, i + 1) > 0.1 || (!pi_ && i == mx - 2)) VX(dp, k, j, i) = 0;
                if (j == 0 || S(nv, k, j, i) + S(nv, k, j + 1, i) > 0.1 || (!pj && j == my - 2)) VY(dp, k, j, i) = 0;
                if (k == 0 || S(nv, k, j, i) + S(nv, k + 1, j, i) > 0.1 || (!pk && k == mz - 2)) VZ(dp, k, j, i) = 0;
            }
}

/* ------------------------------------------------------------------------ */
/* InflowFlux (bcs.c:247-626), uniform inlet (-inlet 1) on the k- face       */
/* ------------------------------------------------------------------------ */

/* Calc_Inlet_Area's inletArea (bcs.c:201-234): |zet| of the k = 0 faces
 * whose first cell is fluid. */
double orc_inlet_area(const orc_params *p, const orc_metrics *m)
{
    double a = 0;
    for (int j = 1; j < p->my - 1; j++)
        for (int i = 1; i < p->mx - 1; i++) {
            double k_area = sqrt(VX(m->zet, 0, j, i) * VX(m->zet, 0, j, i) + VY(m->zet, 0, j, i) * VY(m->zet, 0, j, i) +
                                 VZ(m->zet, 0, j, i) * VZ(m->zet, 0, j, i));
            if (S(m->nvert, 1, j, i) < 0.1) a += k_area;
        }
    return a;
}

/* InletInterpolation (bcs.c:636-681) returns the constant Flux_in of main.c:13
 * before its table code is reached. */
#define ORC_FLUX_IN 4.104388e-04

/* uin of bcs.c:308-371 for the k- inlet cell whose first cell has face
 * centre (xc, yc) and radius r; *ok = 0 for a profile the oracle does not
 * restate (11 adds time-seeded noise; 13, 16, 20, 101 leave uin unset). */
static double inlet_profile(const orc_params *p, double inlet_area, double xc, double yc, double r, int *ok)
{
    (void)xc;
    *ok = 1;
    switch (p->inletprofile) {
    case 0:
    case 3: return ORC_FLUX_IN;
    case 1:
    case 2: return p->inlet_flux < 0 ? 1. : p->inlet_flux / inlet_area;
    case 10: {
        const double delta = 0.45263, a = 5.99;
        if (yc >= delta) return 1.0;
        if (yc <= 0) return 0.0;
        return pow(yc / delta, 1. / a);
    }
    case 12: return 2 * (1 - pow(r / 0.5, 2.0));
    case 14: {
        static const double R[11] = {0.000, 0.062, 0.125, 0.188, 0.250, 0.312, 0.354, 0.399, 0.438, 0.469, 0.500};
        static const double W[11] = {1.120, 1.115, 1.110, 1.093, 1.050, 1.010, 0.963, 0.915, 0.825, 0.730, 0.000};
        int ii;
        for (ii = 1; ii < 11; ii++)
            if (R[ii] >= r && R[ii - 1] < r) break;
        if (ii == 11) ii = 10;  /* r = 0 or r > 0.5 (the value is replaced below for r > 0.5) */
        double u = (W[ii] - W[ii - 1]) / (R[ii] - R[ii - 1]) * (r - R[ii - 1]) + W[ii - 1];
        return r > 0.5 ? 0 : u;
    }
    case 15: return r > 0.5 ? 0 : 1;
    case 21: return 1.5 * 1.0 * yc * (2 - yc);
    case 100: return 0;  /* the plane data are used instead */
    default: *ok = 0; return 0;
    }
}

/* InflowFlux (bcs.c:247-626) for the k- inlet (bctype 5) with -inlet 0, 1, 2,
 * 3, 10, 12, 14, 15, 21 and 100 (plane data p->inflow_plane, my x mx
 * Cmpnts of Cartesian velocity) and the pseudo-periodic inlet.  Returns
 * FluxInSum (the previous value when no k- inlet is configured). */
double orc_inflow_flux(const orc_params *p, const orc_metrics *m, double *ucont, double *ucat)
{
    const int mx = p->mx, my = p->my;
    const double *nv = m->nvert, *coor = m->coor;
    const double th = 0.1;
    double FluxIn = 0;
    if (p->bctype[4] == 5) {
        const int plane = p->pseudo_periodic || p->inletprofile == 100;
        if (plane && !p->inflow_plane) { fprintf(stderr, "oracle: inlet plane data missing\n"); abort(); }
        const double area_in = orc_inlet_area(p, m);
        double flux0 = 0, flux1 = 0;
        const int k = 0;
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                double xc = 0, yc = 0;
                if (coor) {
                    xc = (VX(coor, k + 1, j, i) + VX(coor, k + 1, j - 1, i) + VX(coor, k + 1, j, i - 1) + VX(coor, k + 1, j - 1, i - 1)) * 0.25;
                    yc = (VY(coor, k + 1, j, i) + VY(coor, k + 1, j - 1, i) + VY(coor, k + 1, j, i - 1) + VY(coor, k + 1, j - 1, i - 1)) * 0.25;
                }
                const double r = sqrt(xc * xc + yc * yc);
                const double area = sqrt(VX(m->zet, k, j, i) * VX(m->zet, k, j, i) + VY(m->zet, k, j, i) * VY(m->zet, k, j, i) +
                                          VZ(m->zet, k, j, i) * VZ(m->zet, k, j, i));
                int ok;
                const double uin = inlet_profile(p, area_in, xc, yc, r, &ok);
                if (!ok) { fprintf(stderr, "oracle: -inlet %d is not restated\n", p->inletprofile); abort(); }
                if (S(nv, k + 1, j, i) < th) {
                    if (plane) {
                        const double *pl = p->inflow_plane + 3 * ((size_t)j * mx + i);
                        for (int c = 0; c < 3; c++) VC(ucat, c, k, j, i) = pl[c];
                        VZ(ucont, k, j, i) = 0.5 * (VX(ucat, k, j, i) + VX(ucat, k + 1, j, i)) * VX(m->zet, k, j, i) +
                                             0.5 * (VY(ucat, k, j, i) + VY(ucat, k + 1, j, i)) * VY(m->zet, k, j, i) +
                                             0.5 * (VZ(ucat, k, j, i) + VZ(ucat, k + 1, j, i)) * VZ(m->zet, k, j, i);
                    } else {
                        VZ(ucont, k, j, i) = uin * area;
                        VX(ucont, k, j, i) = 0;
                        VY(ucont, k, j, i) = 0;
                        double q[3] = {0, 0, VZ(ucont, k, j, i)}, u[3];
                        cramer(&VX(m->csi, k, j, i), &VX(m->eta, k, j, i), &VX(m->zet, k, j, i), q, u);
                        for (int c = 0; c < 3; c++) VC(ucat, c, k, j, i) = -VC(ucat, c, k + 1, j, i) + 2 * u[c];
                    }
                } else {
                    VZ(ucat, k, j, i) = 0;
                    VZ(ucont, k, j, i) = 0;
                }
                flux0 += VZ(ucont, k, j, i);
                flux1 += VZ(ucont, k, j, i);
            }
        /* bcs.c:478-496: a held flux for the plane / pseudo-periodic / pipe inlets */
        const int held = p->pseudo_periodic || p->inletprofile == 100 || p->inletprofile == 14;
        if (held) flux0 = p->inlet_flux;
        if (held && p->inlet_flux > 0)
            for (int j = 1; j < my - 1; j++)
                for (int i = 1; i < mx - 1; i++) {
                    const double A = sqrt(VX(m->zet, k, j, i) * VX(m->zet, k, j, i) + VY(m->zet, k, j, i) * VY(m->zet, k, j, i) +
                                          VZ(m->zet, k, j, i) * VZ(m->zet, k, j, i));
                    if (S(nv, k + 1, j, i) < th) VZ(ucont, k, j, i) += (flux0 - flux1) * A / area_in;
                }
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++)
                if (S(nv, k + 1, j, i) < th) FluxIn += VZ(ucont, k, j, i);
    } else {
        return p->flux_in_sum;
    }
    return FluxIn;
}

/* pseudo_periodic_BC (bcs.c:29-95): the Cartesian velocity of plane
 * kplane = (6 mz) / 10 as the next inlet plane (my x mx Cmpnts). */
void orc_pseudo_periodic_plane(const orc_params *p, const double *ucat, double *plane)
{
    const int kplane = (p->mz * 6) / 10;
    for (int j = 0; j < p->my; j++)
        for (int i = 0; i < p->mx; i++)
            for (int c = 0; c < 3; c++) plane[3 * ((size_t)j * p->mx + i) + c] = VC(ucat, c, kplane, j, i);
}
