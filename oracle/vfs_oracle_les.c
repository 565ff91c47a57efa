/*
 * vfs_oracle_les.c -- Compute_Smagorinsky_Constant_1 (les.c:75-1141).
 * TEST INFRASTRUCTURE ONLY (see vfs_oracle.h).
 *
 * les == 1: the constant lCs = Cs^2 (les.c:82-85); les == 2: the dynamic
 * (Germano) procedure with Lilly's least-squares contraction in curvilinear
 * form, the 3x3x3 Simpson test filter of integrate_testfilter_simpson
 * (rhs2.c:499-523, testfilter_ik = 0), optional homogeneous averaging
 * (i_homo_filter / j_homo_filter / k_homo_filter, les.c:799-958), and the
 * clipping to [0, max_cs] (les.c:960-980).  The Clark term (clark = 0 in
 * every shipped deck) and les == 3 are not restated.
 */
#include "orc_internal.h"

static const double LES_EPS = 1.e-4; /* les.c:27: this constant shadows the -les_eps option in les.c */
static const double WALL_CS = 0.001; /* les.c:28 */

/* integrate_testfilter_simpson (rhs2.c:499-523): weights 1 and 4 per
 * direction times w; first index outermost. */
static double simpson(double val[3][3][3], double w[3][3][3])
{
    double wsum = 0, valsum = 0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            for (int c = 0; c < 3; c++) {
                double sw = 1.0;
                if (a == 1) sw *= 4.;
                if (b == 1) sw *= 4.;
                if (c == 1) sw *= 4.;
                wsum += sw * w[a][b][c];
                valsum += sw * w[a][b][c] * val[a][b][c];
            }
    return valsum / wsum;
}

/* get_weight (les.c:30-73): the volume 1/aj of each of the 27 cells, 0 where
 * nvert > nv.  Ghost cells keep their weight (the reference's boundary
 * rules there are commented out).  [R][Q][P] = [k][j][i] offsets + 1. */
static void get_weight(const orc_params *p, const orc_metrics *m, int i, int j, int k, double nv,
                       double w[3][3][3])
{
    for (int pp = -1; pp <= 1; pp++)
        for (int q = -1; q <= 1; q++)
            for (int r = -1; r <= 1; r++) {
                int K = k + r, J = j + q, I = i + pp;
                if (S(m->nvert, K, J, I) > nv) w[r + 1][q + 1][pp + 1] = 0;
                else w[r + 1][q + 1][pp + 1] = 1. / S(m->aj, K, J, I);
            }
}

/* du/dx_r of component c at cell (k,j,i) (Compute_du_center, k-omega.c:313-431,
 * then Compute_du_dxyz, k-omega.c:605-618): g[c][r] */
static void grad_cell(const orc_params *p, const orc_metrics *m, const double *u, int k, int j, int i,
                      double g[3][3])
{
    size_t q = IX(p, k, j, i);
    const double *cs = &m->csi[3 * q], *et = &m->eta[3 * q], *ze = &m->zet[3 * q];
    double d[3][3];
    orc_du_center(p, m->nvert, u, k, j, i, d);
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) g[c][r] = (d[c][0] * cs[r] + d[c][1] * et[r] + d[c][2] * ze[r]) * m->aj[q];
}

static void strain(const double g[3][3], double Sm[3][3])
{
    Sm[0][0] = 0.5 * (g[0][0] + g[0][0]); Sm[0][1] = 0.5 * (g[0][1] + g[1][0]); Sm[0][2] = 0.5 * (g[0][2] + g[2][0]);
    Sm[1][0] = Sm[0][1]; Sm[1][1] = 0.5 * (g[1][1] + g[1][1]); Sm[1][2] = 0.5 * (g[1][2] + g[2][1]);
    Sm[2][0] = Sm[0][2]; Sm[2][1] = Sm[1][2]; Sm[2][2] = 0.5 * (g[2][2] + g[2][2]);
}

static double strain_abs(double Sm[3][3])
{
    return sqrt(2.0 * (Sm[0][0] * Sm[0][0] + Sm[0][1] * Sm[0][1] + Sm[0][2] * Sm[0][2] + Sm[1][0] * Sm[1][0] +
                       Sm[1][1] * Sm[1][1] + Sm[1][2] * Sm[1][2] + Sm[2][0] * Sm[2][0] + Sm[2][1] * Sm[2][1] +
                       Sm[2][2] * Sm[2][2]));
}

/* the reference's explicit ghost copies for periodic directions
 * (les.c:262-292, 702-731, 1092-1118): f[ghost] = f[source], nc components */
static void ghost_copy(const orc_params *p, double *f, int nc) { orc_periodic_fix(p, f, nc); }

/* homogeneous averages (les.c:799-958): sums over fluid cells (nvert < 0.1)
 * of the interior, in the reference's loop order, divided by the count */
static void homo_average(const orc_params *p, const orc_metrics *m, const double *LM, const double *MM,
                         double *cs)
{
    const int mx = p->mx, my = p->my, mz = p->mz;
    const double *nv = m->nvert;
    if (p->i_homo && p->k_homo) {
        for (int j = 1; j < my - 1; j++) {
            double lm = 0, mm = 0;
            int cnt = 0;
            /* les.c:815-824 sums every plane in (k, j, i) order into slot j */
            for (int k = 1; k < mz - 1; k++)
                for (int i = 1; i < mx - 1; i++)
                    if (S(nv, k, j, i) < 0.1) { lm += S(LM, k, j, i); mm += S(MM, k, j, i); cnt++; }
            if (cnt > 0) { lm /= (double)cnt; mm /= (double)cnt; }
            for (int k = 1; k < mz - 1; k++)
                for (int i = 1; i < mx - 1; i++) S(cs, k, j, i) = 0.5 * lm / (mm + LES_EPS);
        }
        return;
    }
    /* one direction (les.c:869-955): a line average for every position in the
     * other two directions, assigned to the fluid cells of that line */
    const int dir = p->i_homo ? 0 : p->j_homo ? 1 : 2;
    const int len = dir == 0 ? mx : dir == 1 ? my : mz;
    const int na = dir == 2 ? my : mz, nb = dir == 0 ? my : mx;
    for (int a = 1; a < na - 1; a++)
        for (int b = 1; b < nb - 1; b++) {
            double lm = 0, mm = 0;
            int cnt = 0;
            for (int t = 1; t < len - 1; t++) {
                int k = dir == 2 ? t : a, j = dir == 0 ? b : dir == 1 ? t : a, i = dir == 0 ? t : b;
                if (S(nv, k, j, i) < 0.1) { lm += S(LM, k, j, i); mm += S(MM, k, j, i); cnt++; }
            }
            if (cnt > 0) { double N = (double)cnt; lm /= N; mm /= N; }
            for (int t = 1; t < len - 1; t++) {
                int k = dir == 2 ? t : a, j = dir == 0 ? b : dir == 1 ? t : a, i = dir == 0 ? t : b;
                if (S(nv, k, j, i) < 0.1) S(cs, k, j, i) = 0.5 * lm / (mm + LES_EPS);
            }
        }
}

void orc_smagorinsky_constant(const orc_params *p, const orc_metrics *m, const double *ucat, int ti, int tistart,
                              int rstart, double *cs)
{
    const int mx = p->mx, my = p->my, mz = p->mz;
    const size_t n = npts(p);
    const double *nv = m->nvert;
    /* les.c:77-85 */
    if (ti < 2 && tistart == 0 && !rstart) { memset(cs, 0, n * sizeof(double)); return; }
    if (p->les == 1) { for (size_t q = 0; q < n; q++) cs[q] = p->cs2; return; }

    /* les.c:258-261: gradients, |S| and the test-filtered velocity */
    double *Ax = dalloc(3 * n), *Ay = dalloc(3 * n), *Az = dalloc(3 * n), *Sabs = dalloc(n), *uf = dalloc(3 * n);
    double *LM = dalloc(n), *MM = dalloc(n);
#pragma omp parallel for schedule(static)
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                if (S(nv, k, j, i) > 1.1) continue;
                double g[3][3], Sm[3][3];
                grad_cell(p, m, ucat, k, j, i, g);
                strain(g, Sm);
                size_t q = IX(p, k, j, i);
                Sabs[q] = strain_abs(Sm);
                for (int r = 0; r < 3; r++) { Ax[3 * q + r] = g[0][r]; Ay[3 * q + r] = g[1][r]; Az[3 * q + r] = g[2][r]; }
                double w[3][3][3], u[3][3][3][3];
                get_weight(p, m, i, j, k, 0.1, w);
                for (int pp = -1; pp <= 1; pp++)
                    for (int qq = -1; qq <= 1; qq++)
                        for (int r = -1; r <= 1; r++)
                            for (int c = 0; c < 3; c++) u[c][r + 1][qq + 1][pp + 1] = VC(ucat, c, k + r, j + qq, i + pp);
                for (int c = 0; c < 3; c++) uf[3 * q + c] = simpson(u[c], w);
            }
    ghost_copy(p, Ax, 3); ghost_copy(p, Ay, 3); ghost_copy(p, Az, 3); ghost_copy(p, Sabs, 1); ghost_copy(p, uf, 3);

    /* les.c:294-697: L_ij, M_ij and their contractions LM, MM per cell */
    static const double coef[3][3][3] = {{{0.125, 0.250, 0.125}, {0.250, 0.500, 0.250}, {0.125, 0.250, 0.125}},
                                         {{0.250, 0.500, 0.250}, {0.500, 1.000, 0.500}, {0.250, 0.500, 0.250}},
                                         {{0.125, 0.250, 0.125}, {0.250, 0.500, 0.250}, {0.125, 0.250, 0.125}}};
#pragma omp parallel for schedule(static)
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t q0 = IX(p, k, j, i);
                if (S(nv, k, j, i) > 1.1) { LM[q0] = MM[q0] = 0; continue; }
                const double *c0 = &m->csi[3 * q0], *e0 = &m->eta[3 * q0], *z0 = &m->zet[3 * q0];
                double w[3][3][3];
                get_weight(p, m, i, j, k, 0.1, w);
                /* the stencil products: Uu[a][b] = (contravariant a) * u_b, SS[a][b] = S_ab |S| */
                double Uu[3][3][3][3][3], SS[3][3][3][3][3];
                for (int pp = -1; pp <= 1; pp++)
                    for (int qq = -1; qq <= 1; qq++)
                        for (int r = -1; r <= 1; r++) {
                            int R = r + 1, Q = qq + 1, P = pp + 1;
                            size_t s = IX(p, k + r, j + qq, i + pp);
                            double u = ucat[3 * s], v = ucat[3 * s + 1], ww = ucat[3 * s + 2];
                            const double *cs = &m->csi[3 * s], *et = &m->eta[3 * s], *ze = &m->zet[3 * s];
                            double U[3];
                            U[0] = u * cs[0] + v * cs[1] + ww * cs[2];
                            U[1] = u * et[0] + v * et[1] + ww * et[2];
                            U[2] = u * ze[0] + v * ze[1] + ww * ze[2];
                            double uc[3] = {u, v, ww};
                            for (int a = 0; a < 3; a++)
                                for (int b = 0; b < 3; b++) Uu[a][b][R][Q][P] = U[a] * uc[b];
                            double g[3][3], Sm[3][3];
                            for (int r2 = 0; r2 < 3; r2++) { g[0][r2] = Ax[3 * s + r2]; g[1][r2] = Ay[3 * s + r2]; g[2][r2] = Az[3 * s + r2]; }
                            strain(g, Sm);
                            for (int a = 0; a < 3; a++)
                                for (int b = 0; b < 3; b++) SS[a][b][R][Q][P] = Sm[a][b] * Sabs[s];
                        }
                double sum_weight = 0;
                for (int pp = -1; pp <= 1; pp++)
                    for (int qq = -1; qq <= 1; qq++)
                        for (int r = -1; r <= 1; r++) sum_weight += w[r + 1][qq + 1][pp + 1] * coef[r + 1][qq + 1][pp + 1];
                double filter = pow(1. / m->aj[q0], 1. / 3.);
                double test_filter = pow(sum_weight, 1. / 3.);
                double fu[3] = {uf[3 * q0], uf[3 * q0 + 1], uf[3 * q0 + 2]};
                double fU[3];
                fU[0] = fu[0] * c0[0] + fu[1] * c0[1] + fu[2] * c0[2];
                fU[1] = fu[0] * e0[0] + fu[1] * e0[1] + fu[2] * e0[2];
                fU[2] = fu[0] * z0[0] + fu[1] * z0[1] + fu[2] * z0[2];
                double gh[3][3], Sh[3][3];
                grad_cell(p, m, uf, k, j, i, gh);  /* gradient of the filtered velocity */
                strain(gh, Sh);
                double S_hat = strain_abs(Sh);
                double L[3][3], SSh[3][3], Mc[3][3], M[3][3];
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) L[a][b] = simpson(Uu[a][b], w) - fU[a] * fu[b];
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) SSh[a][b] = simpson(SS[a][b], w);
                double gm[3][3] = {{c0[0], c0[1], c0[2]}, {e0[0], e0[1], e0[2]}, {z0[0], z0[1], z0[2]}}, gc[3][3], G[3][3];
                orc_covariant(gm, gc);
                /* les.c:600-606: G[a][b] = sum_r gc[r][a] gc[r][b] */
                G[0][0] = gc[0][0] * gc[0][0] + gc[1][0] * gc[1][0] + gc[2][0] * gc[2][0];
                G[1][1] = gc[0][1] * gc[0][1] + gc[1][1] * gc[1][1] + gc[2][1] * gc[2][1];
                G[2][2] = gc[0][2] * gc[0][2] + gc[1][2] * gc[1][2] + gc[2][2] * gc[2][2];
                G[0][1] = G[1][0] = gc[0][1] * gc[0][0] + gc[1][1] * gc[1][0] + gc[2][1] * gc[2][0];
                G[0][2] = G[2][0] = gc[0][2] * gc[0][0] + gc[1][2] * gc[1][0] + gc[2][2] * gc[2][0];
                G[1][2] = G[2][1] = gc[0][1] * gc[0][2] + gc[1][1] * gc[1][2] + gc[2][1] * gc[2][2];
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++)
                        Mc[a][b] = -pow(test_filter, 2.) * S_hat * Sh[a][b] + pow(filter, 2.) * SSh[a][b];
                for (int a = 0; a < 3; a++) {
                    M[a][0] = Mc[a][0] * c0[0] + Mc[a][1] * c0[1] + Mc[a][2] * c0[2];
                    M[a][1] = Mc[a][0] * e0[0] + Mc[a][1] * e0[1] + Mc[a][2] * e0[2];
                    M[a][2] = Mc[a][0] * z0[0] + Mc[a][1] * z0[1] + Mc[a][2] * z0[2];
                }
                double num = 0, denom = 0;
                for (int qq = 0; qq < 3; qq++)
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) num += L[b][a] * M[a][qq] * G[b][qq];
                for (int mm = 0; mm < 3; mm++)
                    for (int nn = 0; nn < 3; nn++)
                        for (int l = 0; l < 3; l++) denom += M[nn][mm] * M[nn][l] * G[mm][l];
                LM[q0] = num;
                MM[q0] = denom;
            }
    ghost_copy(p, LM, 1); ghost_copy(p, MM, 1);

    /* les.c:733-797: filtered (or local, with a homogeneous filter) LM / MM */
    const int homo = p->i_homo || p->j_homo || p->k_homo;
#pragma omp parallel for schedule(static)
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t q0 = IX(p, k, j, i);
                if (S(nv, k, j, i) > 1.1) { cs[q0] = 0; continue; }
                double lm_avg, mm_avg;
                if (homo) {
                    lm_avg = LM[q0];
                    mm_avg = MM[q0];
                } else {
                    double w[3][3][3], L0[3][3][3], M0[3][3][3];
                    for (int a = -1; a <= 1; a++)
                        for (int b = -1; b <= 1; b++)
                            for (int c = -1; c <= 1; c++) {
                                int R = c + 1, Q = b + 1, P = a + 1;
                                int K = k + c, J = j + b, I = i + a;
                                w[R][Q][P] = 1. / S(m->aj, K, J, I);
                                if (S(nv, K, J, I) > 1.1) w[R][Q][P] = 0;
                                if (p->per_i) { if (I == 0) I = mx - 2; else if (I == mx - 1) I = 1; }
                                else if (I == 0 || I == mx - 1) w[R][Q][P] = 0;
                                if (p->per_j) { if (J == 0) J = my - 2; else if (J == my - 1) J = 1; }
                                /* les.c:763: tests the centre's j (never my-1 here), so only J == 0 drops */
                                else if (J == 0 || j == my - 1) w[R][Q][P] = 0;
                                if (p->per_k) { if (K == 0) K = mz - 2; else if (K == mz - 1) K = 1; }
                                else if (K == 0 || K == mz - 1) w[R][Q][P] = 0;
                                L0[R][Q][P] = S(LM, K, J, I);
                                M0[R][Q][P] = S(MM, K, J, I);
                            }
                    lm_avg = simpson(L0, w);
                    mm_avg = simpson(M0, w);
                }
                double C = 0.5 * lm_avg / (mm_avg + LES_EPS);
                cs[q0] = C < 0 ? 0 : C;  /* PetscMax(C, 0) */
            }
    if (homo) homo_average(p, m, LM, MM, cs);

    /* les.c:960-980 */
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                size_t q0 = IX(p, k, j, i);
                if (nv[q0] > 1.1 || k == 0 || k == mz - 1 || j == 0 || j == my - 1 || i == 0 || i == mx - 1) {
                    cs[q0] = 0;
                } else {
                    /* PetscMax(a, b) = a < b ? b : a, PetscMin(a, b) = a < b ? a : b */
                    if (nv[q0] > 0.1 && nv[q0] < 1.1) cs[q0] = WALL_CS < cs[q0] ? cs[q0] : WALL_CS;
                    double c = cs[q0] < 0 ? 0 : cs[q0];
                    cs[q0] = c < p->max_cs ? c : p->max_cs;
                }
            }
    ghost_copy(p, cs, 1); /* les.c:1092-1118 */
    free(Ax); free(Ay); free(Az); free(Sabs); free(uf); free(LM); free(MM);
}
