/*
 * vfs_oracle_rhs.c -- momentum right-hand side and explicit RK4 driver.
 * TEST INFRASTRUCTURE ONLY (see vfs_oracle.h).
 *
 * Restates Formfunction_2 (momentum.c:454-2013), Compute_du_i/j/k
 * (k-omega.c:56-311), RungeKutta (rhs.c:1586-1708) and Calc_Minimum_dt
 * (momentum.c:110-201) for the central schemes without level set, RANS,
 * skew-symmetric or Clark terms and without immersed boundaries.
 */
#include "orc_internal.h"

static const double SOLIDF = 0.1; /* k-omega.c:19 */

/* Compute_du_i, k-omega.c:56-116.  d[c][0..2] = d u_c / d(csi, eta, zeta). */
static void du_i(const orc_params *p, const double *nv, const double *u, int k, int j, int i,
                 double d[3][3])
{
    for (int c = 0; c < 3; c++) {
        d[c][0] = VC(u, c, k, j, i + 1) - VC(u, c, k, j, i);
        if (S(nv, k, j + 1, i) > SOLIDF || S(nv, k, j + 1, i + 1) > SOLIDF)
            d[c][1] = (VC(u, c, k, j, i + 1) + VC(u, c, k, j, i) - VC(u, c, k, j - 1, i + 1) - VC(u, c, k, j - 1, i)) * 0.5;
        else if (S(nv, k, j - 1, i) > SOLIDF || S(nv, k, j - 1, i + 1) > SOLIDF)
            d[c][1] = (VC(u, c, k, j + 1, i + 1) + VC(u, c, k, j + 1, i) - VC(u, c, k, j, i + 1) - VC(u, c, k, j, i)) * 0.5;
        else
            d[c][1] = (VC(u, c, k, j + 1, i + 1) + VC(u, c, k, j + 1, i) - VC(u, c, k, j - 1, i + 1) - VC(u, c, k, j - 1, i)) * 0.25;
        if (S(nv, k + 1, j, i) > SOLIDF || S(nv, k + 1, j, i + 1) > SOLIDF)
            d[c][2] = (VC(u, c, k, j, i + 1) + VC(u, c, k, j, i) - VC(u, c, k - 1, j, i + 1) - VC(u, c, k - 1, j, i)) * 0.5;
        else if (S(nv, k - 1, j, i) > SOLIDF || S(nv, k - 1, j, i + 1) > SOLIDF)
            d[c][2] = (VC(u, c, k + 1, j, i + 1) + VC(u, c, k + 1, j, i) - VC(u, c, k, j, i + 1) - VC(u, c, k, j, i)) * 0.5;
        else
            d[c][2] = (VC(u, c, k + 1, j, i + 1) + VC(u, c, k + 1, j, i) - VC(u, c, k - 1, j, i + 1) - VC(u, c, k - 1, j, i)) * 0.25;
    }
}

/* Compute_du_j, k-omega.c:141-200 */
static void du_j(const orc_params *p, const double *nv, const double *u, int k, int j, int i,
                 double d[3][3])
{
    for (int c = 0; c < 3; c++) {
        if (S(nv, k, j, i + 1) > SOLIDF || S(nv, k, j + 1, i + 1) > SOLIDF)
            d[c][0] = (VC(u, c, k, j + 1, i) + VC(u, c, k, j, i) - VC(u, c, k, j + 1, i - 1) - VC(u, c, k, j, i - 1)) * 0.5;
        else if (S(nv, k, j, i - 1) > SOLIDF || S(nv, k, j + 1, i - 1) > SOLIDF)
            d[c][0] = (VC(u, c, k, j + 1, i + 1) + VC(u, c, k, j, i + 1) - VC(u, c, k, j + 1, i) - VC(u, c, k, j, i)) * 0.5;
        else
            d[c][0] = (VC(u, c, k, j + 1, i + 1) + VC(u, c, k, j, i + 1) - VC(u, c, k, j + 1, i - 1) - VC(u, c, k, j, i - 1)) * 0.25;
        d[c][1] = VC(u, c, k, j + 1, i) - VC(u, c, k, j, i);
        if (S(nv, k + 1, j, i) > SOLIDF || S(nv, k + 1, j + 1, i) > SOLIDF)
            d[c][2] = (VC(u, c, k, j + 1, i) + VC(u, c, k, j, i) - VC(u, c, k - 1, j + 1, i) - VC(u, c, k - 1, j, i)) * 0.5;
        else if (S(nv, k - 1, j, i) > SOLIDF || S(nv, k - 1, j + 1, i) > SOLIDF)
            d[c][2] = (VC(u, c, k + 1, j + 1, i) + VC(u, c, k + 1, j, i) - VC(u, c, k, j + 1, i) - VC(u, c, k, j, i)) * 0.5;
        else
            d[c][2] = (VC(u, c, k + 1, j + 1, i) + VC(u, c, k + 1, j, i) - VC(u, c, k - 1, j + 1, i) - VC(u, c, k - 1, j, i)) * 0.25;
    }
}

/* Compute_du_k, k-omega.c:227-288 */
static void du_k(const orc_params *p, const double *nv, const double *u, int k, int j, int i,
                 double d[3][3])
{
    for (int c = 0; c < 3; c++) {
        if (S(nv, k, j, i + 1) > SOLIDF || S(nv, k + 1, j, i + 1) > SOLIDF)
            d[c][0] = (VC(u, c, k + 1, j, i) + VC(u, c, k, j, i) - VC(u, c, k + 1, j, i - 1) - VC(u, c, k, j, i - 1)) * 0.5;
        else if (S(nv, k, j, i - 1) > SOLIDF || S(nv, k + 1, j, i - 1) > SOLIDF)
            d[c][0] = (VC(u, c, k + 1, j, i + 1) + VC(u, c, k, j, i + 1) - VC(u, c, k + 1, j, i) - VC(u, c, k, j, i)) * 0.5;
        else
            d[c][0] = (VC(u, c, k + 1, j, i + 1) + VC(u, c, k, j, i + 1) - VC(u, c, k + 1, j, i - 1) - VC(u, c, k, j, i - 1)) * 0.25;
        if (S(nv, k, j + 1, i) > SOLIDF || S(nv, k + 1, j + 1, i) > SOLIDF)
            d[c][1] = (VC(u, c, k + 1, j, i) + VC(u, c, k, j, i) - VC(u, c, k + 1, j - 1, i) - VC(u, c, k, j - 1, i)) * 0.5;
        else if (S(nv, k, j - 1, i) > SOLIDF || S(nv, k + 1, j - 1, i) > SOLIDF)
            d[c][1] = (VC(u, c, k + 1, j + 1, i) + VC(u, c, k, j + 1, i) - VC(u, c, k + 1, j, i) - VC(u, c, k, j, i)) * 0.5;
        else
            d[c][1] = (VC(u, c, k + 1, j + 1, i) + VC(u, c, k, j + 1, i) - VC(u, c, k + 1, j - 1, i) - VC(u, c, k, j - 1, i)) * 0.25;
        d[c][2] = VC(u, c, k + 1, j, i) - VC(u, c, k, j, i);
    }
}

/* Viscous face flux of one face (momentum.c:682-696, 863-902 and twins).
 * n = the face's own metric row (csi for i faces, eta for j, zet for k);
 * g1..g3 = n . (csi, eta, zet); the transpose term uses r_{cr}. */
static void visc_face(const double cs[3], const double et[3], const double ze[3], const double nrm[3],
                      double ajc, double d[3][3], double nu, double nu_t, int les, int laplacian,
                      double out[3])
{
    double g11 = cs[0] * nrm[0] + cs[1] * nrm[1] + cs[2] * nrm[2];
    double g21 = et[0] * nrm[0] + et[1] * nrm[1] + et[2] * nrm[2];
    double g31 = ze[0] * nrm[0] + ze[1] * nrm[1] + ze[2] * nrm[2];
    /* r[c][q] = d u_c / d x_q * J (momentum.c:686-696): r11=r[0][0], r21=r[1][0], ... */
    double r[3][3];
    for (int c = 0; c < 3; c++)
        for (int q = 0; q < 3; q++) r[c][q] = d[c][0] * cs[q] + d[c][1] * et[q] + d[c][2] * ze[q];
    for (int c = 0; c < 3; c++) out[c] = 0;
    if (les) {
        /* x: g11*dudc + g21*dude + g31*dudz + r11*n0 + r21*n1 + r31*n2, with
         * r11 = r[0][0], r21 = r[1][0], r31 = r[2][0] (the x column), etc. */
        for (int c = 0; c < 3; c++)
            out[c] = (g11 * d[c][0] + g21 * d[c][1] + g31 * d[c][2] + r[0][c] * nrm[0] + r[1][c] * nrm[1] +
                      r[2][c] * nrm[2]) * ajc * (nu_t);
    }
    if (laplacian)
        for (int c = 0; c < 3; c++)
            for (int q = 0; q < 3; q++) r[c][q] = 0.;
    for (int c = 0; c < 3; c++)
        out[c] += (g11 * d[c][0] + g21 * d[c][1] + g31 * d[c][2] + r[0][c] * nrm[0] + r[1][c] * nrm[1] +
                   r[2][c] * nrm[2]) * ajc * (nu);
}

void orc_formfunction2(const orc_params *p, const orc_metrics *m, double *ucont, const double *ucat,
                       const double *nu_t, double *rhs, double scale)
{
    const int mx = p->mx, my = p->my, mz = p->mz;
    const double *nv = m->nvert;
    const size_t n = npts(p);
    double *div[3], *visc[3];
    for (int d = 0; d < 3; d++) { div[d] = dalloc(3 * n); visc[d] = dalloc(3 * n); }
    double *fp = dalloc(3 * n);
    const double nu = 1. / p->ren;

    orc_periodic_fix(p, ucont, 3); /* momentum.c:638-666 */

    /* i faces, momentum.c:668-928 */
#pragma omp parallel for schedule(static)
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 0; i < mx - 1; i++) {
                size_t q = IX(p, k, j, i);
                const double *cs = &m->icsi[3 * q], *et = &m->ieta[3 * q], *ze = &m->izet[3 * q];
                double ajc = m->iaj[q];
                double d[3][3];
                du_i(p, nv, ucat, k, j, i, d);
                int iL = i - 1, iR = i + 2;
                if (i == 0 || i == mx - 2) {
                    if (p->per_i && i == mx - 2) iR = mx + 2;
                    else if (p->per_i && i == 0) iL = -3;
                    else iL = i, iR = i + 1;
                } else if (S(nv, k, j, iL) + S(nv, k, j, iR) > 0.1) iL = i, iR = i + 1;
                if (p->second_order) iL = i, iR = i + 1;
                double uc = VX(ucont, k, j, i);
                double *dv = &div[0][3 * q];
                for (int c = 0; c < 3; c++) {
                    if (p->second_order)
                        dv[c] = -uc * 0.5 * (VC(ucat, c, k, j, i) + VC(ucat, c, k, j, i + 1));
                    else
                        dv[c] = -uc * 0.0625 * (-VC(ucat, c, k, j, iL) + 9. * VC(ucat, c, k, j, i) +
                                                9. * VC(ucat, c, k, j, i + 1) - VC(ucat, c, k, j, iR));
                }
                if (S(nv, k, j, i) + S(nv, k, j, i + 1) > 0.1) dv[0] = dv[1] = dv[2] = 0;
                double nut = 0;
                if (p->les) {
                    if ((i == 0 && !p->per_i) || S(nv, k, j, i) > 0.1) nut = S(nu_t, k, j, i + 1);
                    else if ((i == mx - 2 && !p->per_i) || S(nv, k, j, i + 1) > 0.1) nut = S(nu_t, k, j, i);
                    else nut = 0.5 * (S(nu_t, k, j, i) + S(nu_t, k, j, i + 1));
                }
                visc_face(cs, et, ze, cs, ajc, d, nu, nut, p->les, p->laplacian, &visc[0][3 * q]);
            }
    /* j faces, momentum.c:930-1211 */
#pragma omp parallel for schedule(static)
    for (int k = 1; k < mz - 1; k++)
        for (int j = 0; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t q = IX(p, k, j, i);
                const double *cs = &m->jcsi[3 * q], *et = &m->jeta[3 * q], *ze = &m->jzet[3 * q];
                double ajc = m->jaj[q];
                double d[3][3];
                du_j(p, nv, ucat, k, j, i, d);
                int jL = j - 1, jR = j + 2;
                if (j == 0 || j == my - 2) {
                    if (p->per_j && j == my - 2) jR = my + 2;
                    else if (p->per_j && j == 0) jL = -3;
                    else jL = j, jR = j + 1;
                } else if (S(nv, k, jL, i) + S(nv, k, jR, i) > 0.1) jL = j, jR = j + 1;
                if (p->second_order) jL = j, jR = j + 1;
                double uc = VY(ucont, k, j, i);
                double *dv = &div[1][3 * q];
                for (int c = 0; c < 3; c++) {
                    if (p->second_order)
                        dv[c] = -uc * 0.5 * (VC(ucat, c, k, j, i) + VC(ucat, c, k, j + 1, i));
                    else
                        dv[c] = -uc * 0.0625 * (-VC(ucat, c, k, jL, i) + 9. * VC(ucat, c, k, j, i) +
                                                9. * VC(ucat, c, k, j + 1, i) - VC(ucat, c, k, jR, i));
                }
                if (S(nv, k, j, i) + S(nv, k, j + 1, i) > 0.1) dv[0] = dv[1] = dv[2] = 0;
                double nut = 0;
                if (p->les) {
                    if ((j == 0 && !p->per_j) || S(nv, k, j, i) > 0.1) nut = S(nu_t, k, j + 1, i);
                    else if ((j == my - 2 && !p->per_j) || S(nv, k, j + 1, i) > 0.1) nut = S(nu_t, k, j, i);
                    else nut = 0.5 * (S(nu_t, k, j, i) + S(nu_t, k, j + 1, i));
                }
                visc_face(cs, et, ze, et, ajc, d, nu, nut, p->les, p->laplacian, &visc[1][3 * q]);
            }
    /* k faces, momentum.c:1212-1451 */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t q = IX(p, k, j, i);
                const double *cs = &m->kcsi[3 * q], *et = &m->keta[3 * q], *ze = &m->kzet[3 * q];
                double ajc = m->kaj[q];
                double d[3][3];
                du_k(p, nv, ucat, k, j, i, d);
                int kL = k - 1, kR = k + 2;
                if (k == 0 || k == mz - 2) {
                    if (p->per_k && k == mz - 2) kR = mz + 2;
                    else if (p->per_k && k == 0) kL = -3;
                    else kL = k, kR = k + 1;
                } else if (S(nv, kL, j, i) + S(nv, kR, j, i) > 0.1) kL = k, kR = k + 1;
                if (p->second_order) kL = k, kR = k + 1;
                double uc = VZ(ucont, k, j, i);
                double *dv = &div[2][3 * q];
                for (int c = 0; c < 3; c++) {
                    if (p->second_order)
                        dv[c] = -uc * 0.5 * (VC(ucat, c, k, j, i) + VC(ucat, c, k + 1, j, i));
                    else
                        dv[c] = -uc * 0.0625 * (-VC(ucat, c, kL, j, i) + 9. * VC(ucat, c, k, j, i) +
                                                9. * VC(ucat, c, k + 1, j, i) - VC(ucat, c, kR, j, i));
                }
                if (S(nv, k, j, i) + S(nv, k + 1, j, i) > 0.1) dv[0] = dv[1] = dv[2] = 0;
                double nut = 0;
                if (p->les) {
                    if ((k == 0 && !p->per_k) || S(nv, k, j, i) > 0.1) nut = S(nu_t, k + 1, j, i);
                    else if ((k == mz - 2 && !p->per_k) || S(nv, k + 1, j, i) > 0.1) nut = S(nu_t, k, j, i);
                    else nut = 0.5 * (S(nu_t, k, j, i) + S(nu_t, k + 1, j, i));
                }
                visc_face(cs, et, ze, ze, ajc, d, nu, nut, p->les, p->laplacian, &visc[2][3 * q]);
            }
    /* momentum.c:1504-1546: periodic fix of the six face-flux arrays */
    for (int d = 0; d < 3; d++) { orc_periodic_fix(p, div[d], 3); orc_periodic_fix(p, visc[d], 3); }

    /* cell loop, momentum.c:1548-1678 */
#pragma omp parallel for schedule(static)
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                double dv[3];
                for (int c = 0; c < 3; c++)
                    dv[c] = (VC(div[0], c, k, j, i) - VC(div[0], c, k, j, i - 1) + VC(div[1], c, k, j, i) -
                             VC(div[1], c, k, j - 1, i) + VC(div[2], c, k, j, i) - VC(div[2], c, k - 1, j, i));
                double *f = &fp[3 * IX(p, k, j, i)];
                if (p->second_order) {
                    for (int c = 0; c < 3; c++) f[c] = dv[c];
                } else {
                    int iR = i + 1, iL = i - 2, jR = j + 1, jL = j - 2, kR = k + 1, kL = k - 2;
                    double di = 3., dj = 3., dk = 3.;
                    if (i == 1) { if (p->per_i) iL = -3; else iR = i, iL = i - 1, di = 1.; }
                    else if (i == 2 || i == mx - 3) { if (!p->per_i) iR = i, iL = i - 1, di = 1.; }
                    else if (i == mx - 2) { if (p->per_i) iR = mx + 1; else iR = i, iL = i - 1, di = 1.; }
                    if (j == 1) { if (p->per_j) jL = -3; else jR = j, jL = j - 1, dj = 1.; }
                    else if (j == 2 || j == my - 3) { if (!p->per_j) jR = j, jL = j - 1, dj = 1.; }
                    else if (j == my - 2) { if (p->per_j) jR = my + 1; else jR = j, jL = j - 1, dj = 1.; }
                    if (k == 1) { if (p->per_k) kL = -3; else kR = k, kL = k - 1, dk = 1.; }
                    else if (k == 2 || k == mz - 3) { if (!p->per_k) kR = k, kL = k - 1, dk = 1.; }
                    else if (k == mz - 2) { if (p->per_k) kR = mz + 1; else kR = k, kL = k - 1, dk = 1.; }
                    if (S(nv, k, j, i - 1) + S(nv, k, j, i) + S(nv, k, j, i + 1) > 0.1) iR = i, iL = i - 1, di = 1.;
                    if (S(nv, k, j - 1, i) + S(nv, k, j, i) + S(nv, k, j + 1, i) > 0.1) jR = j, jL = j - 1, dj = 1.;
                    if (S(nv, k - 1, j, i) + S(nv, k, j, i) + S(nv, k + 1, j, i) > 0.1) kR = k, kL = k - 1, dk = 1.;
                    for (int c = 0; c < 3; c++) {
                        double d4 = 0;
                        d4 += (VC(div[0], c, k, j, iR) - VC(div[0], c, k, j, iL)) * (1. / di);
                        d4 += (VC(div[1], c, k, jR, i) - VC(div[1], c, k, jL, i)) * (1. / dj);
                        d4 += (VC(div[2], c, kR, j, i) - VC(div[2], c, kL, j, i)) * (1. / dk);
                        f[c] = 9. / 8. * dv[c] + -1. / 8. * d4;
                    }
                }
                for (int c = 0; c < 3; c++)
                    f[c] += (VC(visc[0], c, k, j, i) - VC(visc[0], c, k, j, i - 1) + VC(visc[1], c, k, j, i) -
                             VC(visc[1], c, k, j - 1, i) + VC(visc[2], c, k, j, i) - VC(visc[2], c, k - 1, j, i)) / 1.0;
            }
    orc_periodic_fix(p, fp, 3); /* momentum.c:1685-1713 */

    /* contravariant projection, momentum.c:1715-1841 */
#pragma omp parallel for schedule(static)
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t q = IX(p, k, j, i);
                size_t qi = IX(p, k, j, i + 1), qj = IX(p, k, j + 1, i), qk = IX(p, k + 1, j, i);
                const double *f0 = &fp[3 * q], *fi = &fp[3 * qi], *fj = &fp[3 * qj], *fk = &fp[3 * qk];
                const double *c0 = &m->csi[3 * q], *c1 = &m->csi[3 * qi];
                const double *e0 = &m->eta[3 * q], *e1 = &m->eta[3 * qj];
                const double *z0 = &m->zet[3 * q], *z1 = &m->zet[3 * qk];
                double *r = &rhs[3 * q];
                r[0] += scale * (0.5 * (c0[0] * f0[0] + c0[1] * f0[1] + c0[2] * f0[2]) +
                                 0.5 * (c1[0] * fi[0] + c1[1] * fi[1] + c1[2] * fi[2])) * m->iaj[q];
                r[1] += scale * (0.5 * (e0[0] * f0[0] + e0[1] * f0[1] + e0[2] * f0[2]) +
                                 0.5 * (e1[0] * fj[0] + e1[1] * fj[1] + e1[2] * fj[2])) * m->jaj[q];
                r[2] += scale * (0.5 * (z0[0] * f0[0] + z0[1] * f0[1] + z0[2] * f0[2]) +
                                 0.5 * (z1[0] * fk[0] + z1[1] * fk[1] + z1[2] * fk[2])) * m->kaj[q];
                if (S(nv, k, j, i) + S(nv, k, j, i + 1) > 0.1 || (!p->per_i && i == mx - 2)) r[0] = 0;
                if (S(nv, k, j, i) + S(nv, k, j + 1, i) > 0.1 || (!p->per_j && j == my - 2)) r[1] = 0;
                if (S(nv, k, j, i) + S(nv, k + 1, j, i) > 0.1 || (!p->per_k && k == mz - 2)) r[2] = 0;
            }
    /* boundary zeroing, momentum.c:1862-1938 */
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++) {
            for (int c = 0; c < 3; c++) VC(rhs, c, k, j, 0) = 0;
            if (!p->per_i) VX(rhs, k, j, mx - 2) = 0;
            for (int c = 0; c < 3; c++) VC(rhs, c, k, j, mx - 1) = 0;
        }
    for (int k = 0; k < mz; k++)
        for (int i = 0; i < mx; i++) {
            for (int c = 0; c < 3; c++) VC(rhs, c, k, 0, i) = 0;
            if (!p->per_j) VY(rhs, k, my - 2, i) = 0;
            for (int c = 0; c < 3; c++) VC(rhs, c, k, my - 1, i) = 0;
        }
    for (int j = 0; j < my; j++)
        for (int i = 0; i < mx; i++) {
            for (int c = 0; c < 3; c++) VC(rhs, c, 0, j, i) = 0;
            if (!p->per_k) VZ(rhs, mz - 2, j, i) = 0;
            for (int c = 0; c < 3; c++) VC(rhs, c, mz - 1, j, i) = 0;
        }
    for (int d = 0; d < 3; d++) { free(div[d]); free(visc[d]); }
    free(fp);
}

/* RungeKutta, rhs.c:1586-1708 (InflowFlux acts only on bctype 5 inlets). */
void orc_runge_kutta(const orc_params *p_in, const orc_metrics *m, double *ucont, double *ucat,
                     const double *ucat_old, const double *nu_t, const double *dp, double *duold)
{
    /* InflowFlux first (rhs.c:1612); its FluxInSum feeds the outlet balance */
    orc_params pq = *p_in;
    const orc_params *p = &pq;
    if (pq.bctype[4] == 5) pq.flux_in_sum = orc_inflow_flux(&pq, m, ucont, ucat);
    const size_t n3 = 3 * npts(p);
    const double alfa[4] = {0.25, 1 / 3., 0.5, 1.};
    double *ucont_o = dalloc(n3), *pucont = dalloc(n3), *rhs = dalloc(n3);
    memcpy(ucont_o, ucont, n3 * sizeof(double));
    orc_form_bcs(p, m, ucont, ucat, ucat_old, 0);
    memcpy(pucont, ucont, n3 * sizeof(double));
    for (int s = 0; s < 4; s++) {
        orc_contra2cart(p, m, ucont, ucat, ucat_old);
        memset(rhs, 0, n3 * sizeof(double));
        orc_formfunction2(p, m, ucont, ucat, nu_t, rhs, 1.0);
#pragma omp parallel for schedule(static)
        for (size_t q = 0; q < n3; q++) rhs[q] += -1 * dp[q];
        double a = alfa[s] * p->dt * p->st;
#pragma omp parallel for schedule(static)
        for (size_t q = 0; q < n3; q++) ucont[q] = a * rhs[q] + pucont[q];
        orc_contra2cart(p, m, ucont, ucat, ucat_old);
    }
    orc_form_bcs(p, m, ucont, ucat, ucat_old, 1);
    if (duold)
        for (size_t q = 0; q < n3; q++) duold[q] = -1. * ucont_o[q] + ucont[q];
    free(ucont_o); free(pucont); free(rhs);
}

/* Calc_Minimum_dt, momentum.c:110-201: the CFL-1 step (only printed there). */
double orc_min_dt(const orc_params *p, const orc_metrics *m, const double *ucont)
{
    double ldt = 1.e7;
    for (int k = 1; k < p->mz - 1; k++)
        for (int j = 1; j < p->my - 1; j++)
            for (int i = 1; i < p->mx - 1; i++) {
                size_t q = IX(p, k, j, i);
                double aj = m->aj[q];
                for (int c = 0; c < 3; c++) {
                    double v = fabs(1. / ucont[3 * q + c] / aj);
                    if (v < ldt) ldt = v;
                }
            }
    return ldt;
}

/* ------------------------------------------------------------------------ */
/* a7b: FormFunction_SNES (momentum.c:2237-2336), time_coeff() == 1          */
/* ------------------------------------------------------------------------ */

/* r = -(u - ucont_o)/dt + 0.5 F(u) + 0.5 rhs_o - dp (+ feul when non-null),
 * with u copied to ucont, the wall fluxes zeroed (momentum.c:2258-2282) and
 * Contra2Cart applied before Formfunction_2.  PETSc VecAXPY order. */
void orc_formfunction_snes(const orc_params *p, const orc_metrics *m, const double *u, const double *ucont_o,
                           const double *rhs_o, const double *dp, const double *feul, double *ucont, double *ucat,
                           double *ucat_old, double *nu_t, double *r)
{
    const size_t n3 = 3 * npts(p);
    const int mx = p->mx, my = p->my, mz = p->mz;
    const int *b = p->bctype;
    memcpy(ucont, u, n3 * sizeof(double));
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                if (i == 0 && b[0] == 1) VX(ucont, k, j, i) = 0;
                if (i == mx - 1 && b[1] == 1) VX(ucont, k, j, i - 1) = 0;
                if (j == 0 && b[2] == 1) VY(ucont, k, j, i) = 0;
                if (j == my - 1 && b[3] == 1) VY(ucont, k, j - 1, i) = 0;
                if (k == 0 && b[4] == 1) VZ(ucont, k, j, i) = 0;
                if (k == mz - 1 && b[5] == 1) VZ(ucont, k - 1, j, i) = 0;
                if (j == my - 1 && b[3] == 2) VY(ucont, k, j - 1, i) = 0;
                if (j == 0 && b[2] == 12) VY(ucont, k, j, i) = 0;
                if (j == my - 1 && b[3] == 12) VY(ucont, k, j - 1, i) = 0;
                if (b[0] == 10 && i == 0 && (j != 0 && j != my - 1 && k != 0 && k != mz - 1)) VX(ucont, k, j, i) = 0;
                if (b[1] == 10 && i == mx - 1 && (j != 0 && j != my - 1 && k != 0 && k != mz - 1)) VX(ucont, k, j, i - 1) = 0;
                if (b[2] == 10 && j == 0 && (i != 0 && i != mx - 1 && k != 0 && k != mz - 1)) VY(ucont, k, j, i) = 0;
                if (abs(b[3]) == 10 && j == my - 1 && (i != 0 && i != mx - 1 && k != 0 && k != mz - 1)) VY(ucont, k, j - 1, i) = 0;
            }
    orc_contra2cart(p, m, ucont, ucat, ucat_old);
    const double dt = p->dt;
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < n3; q++) {
        r[q] = 0.0 + (-1. / dt) * ucont[q];
        r[q] = r[q] + (1. / dt) * ucont_o[q];
    }
    orc_formfunction2(p, m, ucont, ucat, nu_t, r, 0.5);
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < n3; q++) {
        r[q] = r[q] + 0.5 * rhs_o[q];
        r[q] = r[q] + (-1.0) * dp[q];
        if (feul) r[q] = r[q] + feul[q];
    }
}
