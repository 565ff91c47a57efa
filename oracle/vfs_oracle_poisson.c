/*
 * vfs_oracle_poisson.c -- pressure Poisson right-hand side, operator, FP64
 * multigrid-preconditioned FGMRES, pressure update, projection, divergence.
 * TEST INFRASTRUCTURE ONLY (see vfs_oracle.h).
 *
 * The operator is the poisson==1 form of PoissonLHSNew (poisson.c:977-1688):
 * rows only for fluid interior cells, no aj row scaling, with the one-sided
 * cross-derivative switching next to walls and blanked cells.  The solver
 * iterates are not the reference's (HYPRE GMRES + BoomerAMG or PETSc PCMG,
 * neither vendored); only the converged solution is comparable.
 */
#include "orc_internal.h"

/* 19-point stencil slots: offsets (di, dj, dk). */
enum { CP, EP, WP, NP, SP, TP, BP, NE, SE, NW, SW, TN, BN, TS, BS, TE, BE, TW, BW, NSTENCIL };
static const int OFF[NSTENCIL][3] = {
    {0, 0, 0},  {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},  {0, 0, 1},  {0, 0, -1},
    {1, 1, 0},  {1, -1, 0},  {-1, 1, 0}, {-1, -1, 0}, {0, 1, 1},   {0, 1, -1}, {0, -1, 1},
    {0, -1, -1}, {1, 0, 1},  {1, 0, -1}, {-1, 0, 1},  {-1, 0, -1}};

static inline int is_row(const orc_params *p, const orc_metrics *m, int k, int j, int i)
{
    if (i <= 0 || i >= p->mx - 1 || j <= 0 || j >= p->my - 1 || k <= 0 || k >= p->mz - 1) return 0;
    return S(m->nvert, k, j, i) < p->poisson_threshold;
}

/* Periodic image of a neighbour index (gid ghost mapping, poisson.c:284-309). */
static inline int canon(int i, int m, int per)
{
    if (per) {
        if (i <= 0) i += m - 2;
        else if (i >= m - 1) i -= m - 2;
    }
    return i;
}

/* Face coefficient g_ab of a face array at (k,j,i): (n_a . n_b) * face aj,
 * poisson.c:1105-1113 (the g arrays read the periodically fixed metrics). */
static void face_g(const orc_params *p, const orc_metrics *m, int k, int j, int i, double g[9])
{
    size_t q = IX(p, k, j, i);
    const double *ic = &m->icsi[3 * q], *ie = &m->ieta[3 * q], *iz = &m->izet[3 * q];
    const double *jc = &m->jcsi[3 * q], *je = &m->jeta[3 * q], *jz = &m->jzet[3 * q];
    const double *kc = &m->kcsi[3 * q], *ke = &m->keta[3 * q], *kz = &m->kzet[3 * q];
    double iaj = m->iaj[q], jaj = m->jaj[q], kaj = m->kaj[q];
    g[0] = (ic[0] * ic[0] + ic[1] * ic[1] + ic[2] * ic[2]) * iaj;
    g[1] = (ie[0] * ic[0] + ie[1] * ic[1] + ie[2] * ic[2]) * iaj;
    g[2] = (iz[0] * ic[0] + iz[1] * ic[1] + iz[2] * ic[2]) * iaj;
    g[3] = (jc[0] * je[0] + jc[1] * je[1] + jc[2] * je[2]) * jaj;
    g[4] = (je[0] * je[0] + je[1] * je[1] + je[2] * je[2]) * jaj;
    g[5] = (jz[0] * je[0] + jz[1] * je[1] + jz[2] * je[2]) * jaj;
    g[6] = (kc[0] * kz[0] + kc[1] * kz[1] + kc[2] * kz[2]) * kaj;
    g[7] = (ke[0] * kz[0] + ke[1] * kz[1] + ke[2] * kz[2]) * kaj;
    g[8] = (kz[0] * kz[0] + kz[1] * kz[1] + kz[2] * kz[2]) * kaj;
}

/* One row of PoissonLHSNew, poisson.c:1159-1561, poisson==1 (no aj scaling). */
static void assemble_row(const orc_params *p, const orc_metrics *m, int k, int j, int i, double vol[NSTENCIL])
{
    const int mx = p->mx, my = p->my, mz = p->mz;
    const double th = p->poisson_threshold;
    const double *nv = m->nvert;
    const int pi_ = p->per_i, pj = p->per_j, pk = p->per_k;
    const double r = 1.0;
    for (int s = 0; s < NSTENCIL; s++) vol[s] = 0.;
#define NV(kk, jj, ii) S(nv, kk, jj, ii)
    double g[9];
    /* i+1 - i */
    if (NV(k, j, i + 1) < th && (i != mx - 2 || pi_)) {
        face_g(p, m, k, j, i, g);
        vol[CP] -= g[0] / r;
        vol[EP] += g[0] / r;
        if ((j == my - 2 && !pj) || NV(k, j + 1, i) + NV(k, j + 1, i + 1) > th) {
            if (NV(k, j - 1, i) + NV(k, j - 1, i + 1) < th && j != 1) {
                vol[CP] += g[1] * 0.5 / r; vol[EP] += g[1] * 0.5 / r;
                vol[SP] -= g[1] * 0.5 / r; vol[SE] -= g[1] * 0.5 / r;
            }
        } else if ((j == 1 && !pj) || NV(k, j - 1, i) + NV(k, j - 1, i + 1) > th) {
            if (NV(k, j + 1, i) + NV(k, j + 1, i + 1) < th) {
                vol[NP] += g[1] * 0.5 / r; vol[NE] += g[1] * 0.5 / r;
                vol[CP] -= g[1] * 0.5 / r; vol[EP] -= g[1] * 0.5 / r;
            }
        } else {
            vol[NP] += g[1] * 0.25 / r; vol[NE] += g[1] * 0.25 / r;
            vol[SP] -= g[1] * 0.25 / r; vol[SE] -= g[1] * 0.25 / r;
        }
        if ((k == mz - 2 && !pk) || NV(k + 1, j, i) + NV(k + 1, j, i + 1) > th) {
            if (NV(k - 1, j, i) + NV(k - 1, j, i + 1) < th && k != 1) {
                vol[CP] += g[2] * 0.5 / r; vol[EP] += g[2] * 0.5 / r;
                vol[BP] -= g[2] * 0.5 / r; vol[BE] -= g[2] * 0.5 / r;
            }
        } else if ((k == 1 && !pk) || NV(k - 1, j, i) + NV(k - 1, j, i + 1) > th) {
            if (NV(k + 1, j, i) + NV(k + 1, j, i + 1) < th) {
                vol[TP] += g[2] * 0.5 / r; vol[TE] += g[2] * 0.5 / r;
                vol[CP] -= g[2] * 0.5 / r; vol[EP] -= g[2] * 0.5 / r;
            }
        } else {
            vol[TP] += g[2] * 0.25 / r; vol[TE] += g[2] * 0.25 / r;
            vol[BP] -= g[2] * 0.25 / r; vol[BE] -= g[2] * 0.25 / r;
        }
    }
    /* i - i-1 */
    if (NV(k, j, i - 1) < th && (i != 1 || pi_)) {
        face_g(p, m, k, j, i - 1, g);
        vol[CP] -= g[0] / r;
        vol[WP] += g[0] / r;
        if ((j == my - 2 && !pj) || NV(k, j + 1, i) + NV(k, j + 1, i - 1) > th) {
            if (NV(k, j - 1, i) + NV(k, j - 1, i - 1) < th && j != 1) {
                vol[CP] -= g[1] * 0.5 / r; vol[WP] -= g[1] * 0.5 / r;
                vol[SP] += g[1] * 0.5 / r; vol[SW] += g[1] * 0.5 / r;
            }
        } else if ((j == 1 && !pj) || NV(k, j - 1, i) + NV(k, j - 1, i - 1) > th) {
            if (NV(k, j + 1, i) + NV(k, j + 1, i - 1) < th) {
                vol[NP] -= g[1] * 0.5 / r; vol[NW] -= g[1] * 0.5 / r;
                vol[CP] += g[1] * 0.5 / r; vol[WP] += g[1] * 0.5 / r;
            }
        } else {
            vol[NP] -= g[1] * 0.25 / r; vol[NW] -= g[1] * 0.25 / r;
            vol[SP] += g[1] * 0.25 / r; vol[SW] += g[1] * 0.25 / r;
        }
        if ((k == mz - 2 && !pk) || NV(k + 1, j, i) + NV(k + 1, j, i - 1) > th) {
            if (NV(k - 1, j, i) + NV(k - 1, j, i - 1) < th && k != 1) {
                vol[CP] -= g[2] * 0.5 / r; vol[WP] -= g[2] * 0.5 / r;
                vol[BP] += g[2] * 0.5 / r; vol[BW] += g[2] * 0.5 / r;
            }
        } else if ((k == 1 && !pk) || NV(k - 1, j, i) + NV(k - 1, j, i - 1) > th) {
            if (NV(k + 1, j, i) + NV(k + 1, j, i - 1) < th) {
                vol[TP] -= g[2] * 0.5 / r; vol[TW] -= g[2] * 0.5 / r;
                vol[CP] += g[2] * 0.5 / r; vol[WP] += g[2] * 0.5 / r;
            }
        } else {
            vol[TP] -= g[2] * 0.25 / r; vol[TW] -= g[2] * 0.25 / r;
            vol[BP] += g[2] * 0.25 / r; vol[BW] += g[2] * 0.25 / r;
        }
    }
    /* j+1 - j */
    if (NV(k, j + 1, i) < th && (j != my - 2 || pj)) {
        face_g(p, m, k, j, i, g);
        if ((i == mx - 2 && !pi_) || NV(k, j, i + 1) + NV(k, j + 1, i + 1) > th) {
            if (NV(k, j, i - 1) + NV(k, j + 1, i - 1) < th && i != 1) {
                vol[CP] += g[3] * 0.5 / r; vol[NP] += g[3] * 0.5 / r;
                vol[WP] -= g[3] * 0.5 / r; vol[NW] -= g[3] * 0.5 / r;
            }
        } else if ((i == 1 && !pi_) || NV(k, j, i - 1) + NV(k, j + 1, i - 1) > th) {
            if (NV(k, j, i + 1) + NV(k, j + 1, i + 1) < th) {
                vol[EP] += g[3] * 0.5 / r; vol[NE] += g[3] * 0.5 / r;
                vol[CP] -= g[3] * 0.5 / r; vol[NP] -= g[3] * 0.5 / r;
            }
        } else {
            vol[EP] += g[3] * 0.25 / r; vol[NE] += g[3] * 0.25 / r;
            vol[WP] -= g[3] * 0.25 / r; vol[NW] -= g[3] * 0.25 / r;
        }
        vol[CP] -= g[4] / r;
        vol[NP] += g[4] / r;
        if ((k == mz - 2 && !pk) || NV(k + 1, j, i) + NV(k + 1, j + 1, i) > th) {
            if (NV(k - 1, j, i) + NV(k - 1, j + 1, i) < th && k != 1) {
                vol[CP] += g[5] * 0.5 / r; vol[NP] += g[5] * 0.5 / r;
                vol[BP] -= g[5] * 0.5 / r; vol[BN] -= g[5] * 0.5 / r;
            }
        } else if ((k == 1 && !pk) || NV(k - 1, j, i) + NV(k - 1, j + 1, i) > th) {
            if (NV(k + 1, j, i) + NV(k + 1, j + 1, i) < th) {
                vol[TP] += g[5] * 0.5 / r; vol[TN] += g[5] * 0.5 / r;
                vol[CP] -= g[5] * 0.5 / r; vol[NP] -= g[5] * 0.5 / r;
            }
        } else {
            vol[TP] += g[5] * 0.25 / r; vol[TN] += g[5] * 0.25 / r;
            vol[BP] -= g[5] * 0.25 / r; vol[BN] -= g[5] * 0.25 / r;
        }
    }
    /* j - j-1 */
    if (NV(k, j - 1, i) < th && (j != 1 || pj)) {
        face_g(p, m, k, j - 1, i, g);
        if ((i == mx - 2 && !pi_) || NV(k, j, i + 1) + NV(k, j - 1, i + 1) > th) {
            if (NV(k, j, i - 1) + NV(k, j - 1, i - 1) < th && i != 1) {
                vol[CP] -= g[3] * 0.5 / r; vol[SP] -= g[3] * 0.5 / r;
                vol[WP] += g[3] * 0.5 / r; vol[SW] += g[3] * 0.5 / r;
            }
        } else if ((i == 1 && !pi_) || NV(k, j, i - 1) + NV(k, j - 1, i - 1) > th) {
            if (NV(k, j, i + 1) + NV(k, j - 1, i + 1) < th) {
                vol[EP] -= g[3] * 0.5 / r; vol[SE] -= g[3] * 0.5 / r;
                vol[CP] += g[3] * 0.5 / r; vol[SP] += g[3] * 0.5 / r;
            }
        } else {
            vol[EP] -= g[3] * 0.25 / r; vol[SE] -= g[3] * 0.25 / r;
            vol[WP] += g[3] * 0.25 / r; vol[SW] += g[3] * 0.25 / r;
        }
        vol[CP] -= g[4] / r;
        vol[SP] += g[4] / r;
        if ((k == mz - 2 && !pk) || NV(k + 1, j, i) + NV(k + 1, j - 1, i) > th) {
            if (NV(k - 1, j, i) + NV(k - 1, j - 1, i) < th && k != 1) {
                vol[CP] -= g[5] * 0.5 / r; vol[SP] -= g[5] * 0.5 / r;
                vol[BP] += g[5] * 0.5 / r; vol[BS] += g[5] * 0.5 / r;
            }
        } else if ((k == 1 && !pk) || NV(k - 1, j, i) + NV(k - 1, j - 1, i) > th) {
            if (NV(k + 1, j, i) + NV(k + 1, j - 1, i) < th) {
                vol[TP] -= g[5] * 0.5 / r; vol[TS] -= g[5] * 0.5 / r;
                vol[CP] += g[5] * 0.5 / r; vol[SP] += g[5] * 0.5 / r;
            }
        } else {
            vol[TP] -= g[5] * 0.25 / r; vol[TS] -= g[5] * 0.25 / r;
            vol[BP] += g[5] * 0.25 / r; vol[BS] += g[5] * 0.25 / r;
        }
    }
    /* k+1 - k */
    if (NV(k + 1, j, i) < th && (k != mz - 2 || pk)) {
        face_g(p, m, k, j, i, g);
        if ((i == mx - 2 && !pi_) || NV(k, j, i + 1) + NV(k + 1, j, i + 1) > th) {
            if (NV(k, j, i - 1) + NV(k + 1, j, i - 1) < th && i != 1) {
                vol[CP] += g[6] * 0.5 / r; vol[TP] += g[6] * 0.5 / r;
                vol[WP] -= g[6] * 0.5 / r; vol[TW] -= g[6] * 0.5 / r;
            }
        } else if ((i == 1 && !pi_) || NV(k, j, i - 1) + NV(k + 1, j, i - 1) > th) {
            if (NV(k, j, i + 1) + NV(k + 1, j, i + 1) < th) {
                vol[EP] += g[6] * 0.5 / r; vol[TE] += g[6] * 0.5 / r;
                vol[CP] -= g[6] * 0.5 / r; vol[TP] -= g[6] * 0.5 / r;
            }
        } else {
            vol[EP] += g[6] * 0.25 / r; vol[TE] += g[6] * 0.25 / r;
            vol[WP] -= g[6] * 0.25 / r; vol[TW] -= g[6] * 0.25 / r;
        }
        if ((j == my - 2 && !pj) || NV(k, j + 1, i) + NV(k + 1, j + 1, i) > th) {
            if (NV(k, j - 1, i) + NV(k + 1, j - 1, i) < th && j != 1) {
                vol[CP] += g[7] * 0.5 / r; vol[TP] += g[7] * 0.5 / r;
                vol[SP] -= g[7] * 0.5 / r; vol[TS] -= g[7] * 0.5 / r;
            }
        } else if ((j == 1 && !pj) || NV(k, j - 1, i) + NV(k + 1, j - 1, i) > th) {
            if (NV(k, j + 1, i) + NV(k + 1, j + 1, i) < th) {
                vol[NP] += g[7] * 0.5 / r; vol[TN] += g[7] * 0.5 / r;
                vol[CP] -= g[7] * 0.5 / r; vol[TP] -= g[7] * 0.5 / r;
            }
        } else {
            vol[NP] += g[7] * 0.25 / r; vol[TN] += g[7] * 0.25 / r;
            vol[SP] -= g[7] * 0.25 / r; vol[TS] -= g[7] * 0.25 / r;
        }
        vol[CP] -= g[8] / r;
        vol[TP] += g[8] / r;
    }
    /* k - k-1 */
    if (NV(k - 1, j, i) < th && (k != 1 || pk)) {
        face_g(p, m, k - 1, j, i, g);
        if ((i == mx - 2 && !pi_) || NV(k, j, i + 1) + NV(k - 1, j, i + 1) > th) {
            if (NV(k, j, i - 1) + NV(k - 1, j, i - 1) < th && i != 1) {
                vol[CP] -= g[6] * 0.5 / r; vol[BP] -= g[6] * 0.5 / r;
                vol[WP] += g[6] * 0.5 / r; vol[BW] += g[6] * 0.5 / r;
            }
        } else if ((i == 1 && !pi_) || NV(k, j, i - 1) + NV(k - 1, j, i - 1) > th) {
            if (NV(k, j, i + 1) + NV(k - 1, j, i + 1) < th) {
                vol[EP] -= g[6] * 0.5 / r; vol[BE] -= g[6] * 0.5 / r;
                vol[CP] += g[6] * 0.5 / r; vol[BP] += g[6] * 0.5 / r;
            }
        } else {
            vol[EP] -= g[6] * 0.25 / r; vol[BE] -= g[6] * 0.25 / r;
            vol[WP] += g[6] * 0.25 / r; vol[BW] += g[6] * 0.25 / r;
        }
        if ((j == my - 2 && !pj) || NV(k, j + 1, i) + NV(k - 1, j + 1, i) > th) {
            if (NV(k, j - 1, i) + NV(k - 1, j - 1, i) < th && j != 1) {
                vol[CP] -= g[7] * 0.5 / r; vol[BP] -= g[7] * 0.5 / r;
                vol[SP] += g[7] * 0.5 / r; vol[BS] += g[7] * 0.5 / r;
            }
        } else if ((j == 1 && !pj) || NV(k, j - 1, i) + NV(k - 1, j - 1, i) > th) {
            if (NV(k, j + 1, i) + NV(k - 1, j + 1, i) < th) {
                vol[NP] -= g[7] * 0.5 / r; vol[BN] -= g[7] * 0.5 / r;
                vol[CP] += g[7] * 0.5 / r; vol[BP] += g[7] * 0.5 / r;
            }
        } else {
            vol[NP] -= g[7] * 0.25 / r; vol[BN] -= g[7] * 0.25 / r;
            vol[SP] += g[7] * 0.25 / r; vol[BS] += g[7] * 0.25 / r;
        }
        vol[CP] -= g[8] / r;
        vol[BP] += g[8] / r;
    }
#undef NV
}

/* ------------------------------------------------------------------------ */
/* Operator storage for one multigrid level                                  */
/* ------------------------------------------------------------------------ */

typedef struct {
    orc_params p;        /* grid of this level */
    orc_metrics m;       /* metrics of this level (owned for coarse levels) */
    double *coef;        /* NSTENCIL doubles per DA point (0 for non-rows) */
    unsigned char *row;  /* 1 where a fluid row exists */
    int owns_metrics;
} level_t;

static inline size_t nbr(const orc_params *p, int k, int j, int i, int s)
{
    int ii = canon(i + OFF[s][0], p->mx, p->per_i);
    int jj = canon(j + OFF[s][1], p->my, p->per_j);
    int kk = canon(k + OFF[s][2], p->mz, p->per_k);
    return IX(p, kk, jj, ii);
}

static void level_build_operator(level_t *L)
{
    const orc_params *p = &L->p;
    size_t n = npts(p);
    L->coef = dalloc(NSTENCIL * n);
    L->row = (unsigned char *)calloc(n, 1);
#pragma omp parallel for schedule(static)
    for (int k = 1; k < p->mz - 1; k++)
        for (int j = 1; j < p->my - 1; j++)
            for (int i = 1; i < p->mx - 1; i++) {
                if (!is_row(p, &L->m, k, j, i)) continue;
                size_t q = IX(p, k, j, i);
                L->row[q] = 1;
            }
#pragma omp parallel for schedule(static)
    for (int k = 1; k < p->mz - 1; k++)
        for (int j = 1; j < p->my - 1; j++)
            for (int i = 1; i < p->mx - 1; i++) {
                size_t q = IX(p, k, j, i);
                if (!L->row[q]) continue;
                double *c = &L->coef[NSTENCIL * q];
                assemble_row(p, &L->m, k, j, i, c);
                for (int t = 1; t < NSTENCIL; t++)
                    if (!L->row[nbr(p, k, j, i, t)]) c[t] = 0.;
            }
}


static void level_apply(const level_t *L, const double *x, double *y)
{
    const orc_params *p = &L->p;
    memset(y, 0, npts(p) * sizeof(double));
#pragma omp parallel for schedule(static)
    for (int k = 1; k < p->mz - 1; k++)
        for (int j = 1; j < p->my - 1; j++)
            for (int i = 1; i < p->mx - 1; i++) {
                size_t q = IX(p, k, j, i);
                if (!L->row[q]) continue;
                const double *c = &L->coef[NSTENCIL * q];
                double s = 0;
                for (int t = 0; t < NSTENCIL; t++) {
                    size_t nb = nbr(p, k, j, i, t);
                    if (L->row[nb]) s += c[t] * x[nb];
                }
                y[q] = s;
            }
}

void orc_poisson_apply(const orc_params *p, const orc_metrics *m, const double *phi, double *y)
{
    level_t L;
    memset(&L, 0, sizeof(L));
    L.p = *p;
    L.m = *m;
    level_build_operator(&L);
    level_apply(&L, phi, y);
    free(L.coef);
    free(L.row);
}

/* ------------------------------------------------------------------------ */
/* a11: PoissonRHS2_hypre + Remove_Nullspace_Scale                           */
/* ------------------------------------------------------------------------ */

void orc_poisson_rhs(const orc_params *p, const orc_metrics *m, const double *ucont, double *b)
{
    const int mx = p->mx, my = p->my, mz = p->mz;
    memset(b, 0, npts(p) * sizeof(double));
    const double coeff = 1.; /* time_coeff(), poisson.c:23-31 */
#pragma omp parallel for schedule(static)
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                if (S(m->nvert, k, j, i) >= 0.1) continue;
                double val = 0;
                val -= VX(ucont, k, j, i);
                if (i == 1 && p->per_i) val += VX(ucont, k, j, -2);
                else val += VX(ucont, k, j, i - 1);
                val -= VY(ucont, k, j, i);
                if (j == 1 && p->per_j) val += VY(ucont, k, -2, i);
                else val += VY(ucont, k, j - 1, i);
                val -= VZ(ucont, k, j, i);
                if (k == 1 && p->per_k) val += VZ(ucont, -2, j, i);
                else val += VZ(ucont, k - 1, j, i);
                val *= -1.0 / p->dt * p->st * coeff;
                S(b, k, j, i) = val;
            }
    /* Remove_Nullspace_Scale, poisson_hypre.c:118-200 */
    double sum = 0, sum_aj = 0;
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++)
                if (S(m->nvert, k, j, i) < 0.1) { sum += S(b, k, j, i); sum_aj += S(m->aj, k, j, i); }
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++)
                if (S(m->nvert, k, j, i) < 0.1) S(b, k, j, i) += -sum * S(m->aj, k, j, i) / sum_aj;
}

/* Remove_Nullspace on the solution, poisson_hypre.c:202-263 */
void orc_remove_mean(const orc_params *p, const orc_metrics *m, double *phi)
{
    double sum = 0;
    long cnt = 0;
    for (int k = 1; k < p->mz - 1; k++)
        for (int j = 1; j < p->my - 1; j++)
            for (int i = 1; i < p->mx - 1; i++)
                if (S(m->nvert, k, j, i) < 0.1) { sum += S(phi, k, j, i); cnt++; }
    double val = -sum / (double)cnt;
    for (int k = 1; k < p->mz - 1; k++)
        for (int j = 1; j < p->my - 1; j++)
            for (int i = 1; i < p->mx - 1; i++)
                if (S(m->nvert, k, j, i) < 0.1) S(phi, k, j, i) += val;
}

/* ------------------------------------------------------------------------ */
/* Geometric multigrid, the design of the HIP solver restated:               */
/*  - coarse grids take every other node (init.c:402-455), their operators   */
/*    are rediscretised with FormMetrics on the coarse nodes;               */
/*  - restriction sums the 8 child residuals (the rows are flux sums);      */
/*  - prolongation is cell-centred trilinear (weights 27/9/3/1 over 64);     */
/*  - smoother: j-line Gauss-Seidel in 4 colours (i%2, k%2), exact Thomas   */
/*    solves along each wall-normal line.                                    */
/* ------------------------------------------------------------------------ */

typedef struct {
    int nl;
    level_t L[8];
    double *r[8], *x[8], *b[8], *tmp[8];
    double *coor[8];
} mg_t;

static void metrics_alloc(const orc_params *p, orc_metrics *m)
{
    size_t n = npts(p);
    double **v[] = {&m->csi, &m->eta, &m->zet, &m->icsi, &m->ieta, &m->izet, &m->jcsi,
                    &m->jeta, &m->jzet, &m->kcsi, &m->keta, &m->kzet, &m->cent};
    for (size_t q = 0; q < sizeof(v) / sizeof(v[0]); q++) *v[q] = dalloc(3 * n);
    double **s[] = {&m->aj, &m->iaj, &m->jaj, &m->kaj, &m->nvert};
    for (size_t q = 0; q < sizeof(s) / sizeof(s[0]); q++) *s[q] = dalloc(n);
}

static void metrics_free(orc_metrics *m)
{
    double *v[] = {m->csi, m->eta, m->zet, m->icsi, m->ieta, m->izet, m->jcsi, m->jeta,
                   m->jzet, m->kcsi, m->keta, m->kzet, m->cent, m->aj, m->iaj, m->jaj, m->kaj, m->nvert};
    for (size_t q = 0; q < sizeof(v) / sizeof(v[0]); q++) free(v[q]);
}

static void mg_build(mg_t *mg, const orc_params *p, const orc_metrics *m, int levels)
{
    memset(mg, 0, sizeof(*mg));
    mg->nl = levels;
    mg->L[0].p = *p;
    mg->L[0].m = *m;
    level_build_operator(&mg->L[0]);
    const double *fc = m->coor;
    for (int l = 1; l < levels; l++) {
        const orc_params *pf = &mg->L[l - 1].p;
        orc_params pc = *pf;
        int IMf = pf->mx - 1, JMf = pf->my - 1, KMf = pf->mz - 1; /* node counts */
        pc.mx = (IMf + 1) / 2 + 1;
        pc.my = (JMf + 1) / 2 + 1;
        pc.mz = (KMf + 1) / 2 + 1;
        mg->L[l].p = pc;
        metrics_alloc(&pc, &mg->L[l].m);
        mg->L[l].owns_metrics = 1;
        double *cc = dalloc(3 * npts(&pc));
        for (int k = 0; k < pc.mz - 1; k++)
            for (int j = 0; j < pc.my - 1; j++)
                for (int i = 0; i < pc.mx - 1; i++)
                    for (int c = 0; c < 3; c++)
                        cc[3 * (((size_t)k * pc.my + j) * pc.mx + i) + c] =
                            fc[3 * (((size_t)(2 * k) * pf->my + 2 * j) * pf->mx + 2 * i) + c];
        orc_form_metrics(&pc, cc, &mg->L[l].m);
        mg->L[l].m.coor = cc;
        /* a coarse cell is blanked only when all eight children are */
        const orc_params *pp = &pc;
        for (int k = 1; k < pc.mz - 1; k++)
            for (int j = 1; j < pc.my - 1; j++)
                for (int i = 1; i < pc.mx - 1; i++) {
                    double mn = 1e30;
                    for (int dk = 0; dk < 2; dk++)
                        for (int dj = 0; dj < 2; dj++)
                            for (int di = 0; di < 2; di++) {
                                double v = mg->L[l - 1].m.nvert[IX(pf, 2 * k - 1 + dk, 2 * j - 1 + dj, 2 * i - 1 + di)];
                                if (v < mn) mn = v;
                            }
                    mg->L[l].m.nvert[IX(pp, k, j, i)] = mn;
                }
        level_build_operator(&mg->L[l]);
        mg->coor[l] = cc;
        fc = cc;
    }
    for (int l = 0; l < levels; l++) {
        size_t n = npts(&mg->L[l].p);
        mg->r[l] = dalloc(n); mg->x[l] = dalloc(n); mg->b[l] = dalloc(n); mg->tmp[l] = dalloc(n);
    }
}

static void mg_free(mg_t *mg)
{
    for (int l = 0; l < mg->nl; l++) {
        free(mg->L[l].coef); free(mg->L[l].row);
        if (mg->L[l].owns_metrics) metrics_free(&mg->L[l].m);
        free(mg->r[l]); free(mg->x[l]); free(mg->b[l]); free(mg->tmp[l]); free(mg->coor[l]);
    }
}

/* Zebra Gauss-Seidel over wall-normal j lines: the lines (i,k) of colour
 * (i+k)&1 == 0 are relaxed first, then colour 1.  Within one colour every line
 * takes its off-line neighbours from the values at the start of that colour's
 * half-sweep and is solved exactly by the Thomas algorithm. */
static void smooth_lines(const level_t *L, double *x, const double *b)
{
    const orc_params *p = &L->p;
    const int my = p->my;
    size_t n = npts(p);
    double *xin = dalloc(n);
    for (int col = 0; col < 2; col++) {
        memcpy(xin, x, n * sizeof(double));
        /* the lines of one colour are independent: they read xin only */
#pragma omp parallel
        {
        double *cp = dalloc(my), *dp = dalloc(my);
#pragma omp for schedule(static) collapse(2)
        for (int k = 1; k < p->mz - 1; k++)
            for (int i = 1; i < p->mx - 1; i++) {
                if (((i + k) & 1) != col) continue;
                int have_prev = 0;
                for (int j = 1; j < my - 1; j++) {
                    size_t q = IX(p, k, j, i);
                    if (!L->row[q]) { cp[j] = 0; dp[j] = 0; have_prev = 0; continue; }
                    const double *c = &L->coef[NSTENCIL * q];
                    double rhs = b[q];
                    for (int t = 0; t < NSTENCIL; t++) {
                        if (t == CP || t == NP || t == SP) continue;
                        rhs -= c[t] * xin[nbr(p, k, j, i, t)];
                    }
                    double lo = have_prev ? c[SP] : 0.;
                    double den = c[CP] - lo * (have_prev ? cp[j - 1] : 0.);
                    cp[j] = c[NP] / den;
                    dp[j] = (rhs - lo * (have_prev ? dp[j - 1] : 0.)) / den;
                    have_prev = 1;
                }
                double next = 0;
                int next_valid = 0;
                for (int j = my - 2; j >= 1; j--) {
                    size_t q = IX(p, k, j, i);
                    if (!L->row[q]) { x[q] = 0; next_valid = 0; continue; }
                    double xl = dp[j] - (next_valid ? cp[j] * next : 0.);
                    next = xl;
                    next_valid = 1;
                    x[q] = xl;
                }
            }
        free(cp); free(dp);
        }
    }
    free(xin);
}

static void restrict_sum(const level_t *Lf, const double *rf, const level_t *Lc, double *bc)
{
    const orc_params *pc = &Lc->p, *pf = &Lf->p;
    memset(bc, 0, npts(pc) * sizeof(double));
#pragma omp parallel for schedule(static)
    for (int k = 1; k < pc->mz - 1; k++)
        for (int j = 1; j < pc->my - 1; j++)
            for (int i = 1; i < pc->mx - 1; i++) {
                size_t qc = IX(pc, k, j, i);
                if (!Lc->row[qc]) continue;
                double s = 0;
                for (int dk = 0; dk < 2; dk++)
                    for (int dj = 0; dj < 2; dj++)
                        for (int di = 0; di < 2; di++) {
                            size_t qf = IX(pf, 2 * k - 1 + dk, 2 * j - 1 + dj, 2 * i - 1 + di);
                            if (Lf->row[qf]) s += rf[qf];
                        }
                bc[qc] = s;
            }
}

/* coarse index of the neighbouring parent in one direction (wrap / clamp) */
static inline int parent_nb(int I, int d, int mc, int per)
{
    int J = I + d;
    if (J < 1) J = per ? mc - 2 : 1;
    else if (J > mc - 2) J = per ? 1 : mc - 2;
    return J;
}

static void prolong_add(const level_t *Lc, const double *xc, const level_t *Lf, double *xf)
{
    const orc_params *pc = &Lc->p, *pf = &Lf->p;
#pragma omp parallel for schedule(static)
    for (int k = 1; k < pf->mz - 1; k++)
        for (int j = 1; j < pf->my - 1; j++)
            for (int i = 1; i < pf->mx - 1; i++) {
                size_t qf = IX(pf, k, j, i);
                if (!Lf->row[qf]) continue;
                int I = (i + 1) / 2, J = (j + 1) / 2, K = (k + 1) / 2;
                int di = (i & 1) ? -1 : 1, dj = (j & 1) ? -1 : 1, dk = (k & 1) ? -1 : 1;
                int I2 = parent_nb(I, di, pc->mx, pc->per_i);
                int J2 = parent_nb(J, dj, pc->my, pc->per_j);
                int K2 = parent_nb(K, dk, pc->mz, pc->per_k);
                double v = 27. * xc[IX(pc, K, J, I)] + 9. * (xc[IX(pc, K, J, I2)] + xc[IX(pc, K, J2, I)] + xc[IX(pc, K2, J, I)]) +
                           3. * (xc[IX(pc, K, J2, I2)] + xc[IX(pc, K2, J, I2)] + xc[IX(pc, K2, J2, I)]) +
                           xc[IX(pc, K2, J2, I2)];
                xf[qf] += v / 64.;
            }
}

static void residual(const level_t *L, const double *x, const double *b, double *r)
{
    level_apply(L, x, r);
    size_t n = npts(&L->p);
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < n; q++) r[q] = L->row[q] ? b[q] - r[q] : 0.;
}

static void vcycle(mg_t *mg, int l, int pre, int post, int coarse_sweeps)
{
    level_t *L = &mg->L[l];
    if (l == mg->nl - 1 || l == 7) {
        for (int s = 0; s < coarse_sweeps; s++) smooth_lines(L, mg->x[l], mg->b[l]);
        return;
    }
    for (int s = 0; s < pre; s++) smooth_lines(L, mg->x[l], mg->b[l]);
    residual(L, mg->x[l], mg->b[l], mg->r[l]);
    restrict_sum(L, mg->r[l], &mg->L[l + 1], mg->b[l + 1]);
    memset(mg->x[l + 1], 0, npts(&mg->L[l + 1].p) * sizeof(double));
    vcycle(mg, l + 1, pre, post, coarse_sweeps);
    prolong_add(&mg->L[l + 1], mg->x[l + 1], L, mg->x[l]);
    for (int s = 0; s < post; s++) smooth_lines(L, mg->x[l], mg->b[l]);
}

static double dot(const level_t *L, const double *a, const double *b)
{
    /* partial sums over fixed chunks of 4096 points, added in chunk order:
     * the same value whatever the number of threads */
    const size_t CH = 4096, n = npts(&L->p), nch = (n + CH - 1) / CH;
    double *part = dalloc(nch);
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < nch; c++) {
        double s = 0;
        const size_t e = (c + 1) * CH < n ? (c + 1) * CH : n;
        for (size_t q = c * CH; q < e; q++) if (L->row[q]) s += a[q] * b[q];
        part[c] = s;
    }
    double s = 0;
    for (size_t c = 0; c < nch; c++) s += part[c];
    free(part);
    return s;
}

int orc_poisson_solve(const orc_params *p, const orc_metrics *m, const double *b, double *phi,
                      double tol, int restart, int maxit, int levels, double *relres)
{
    mg_t mg;
    if (!m->coor) levels = 1;
    {   /* coarsen only while every interior cell count halves exactly */
        int nl = 1, cx = p->mx - 2, cy = p->my - 2, cz = p->mz - 2;
        while (nl < levels && nl < 8 && cx % 2 == 0 && cy % 2 == 0 && cz % 2 == 0 && cx >= 4 && cy >= 4 && cz >= 4) {
            cx /= 2; cy /= 2; cz /= 2; nl++;
        }
        levels = nl;
    }
    mg_build(&mg, p, m, levels);
    level_t *L = &mg.L[0];
    size_t n = npts(p);
    int mr = restart;
    double **V = (double **)malloc((mr + 1) * sizeof(double *));
    double **Z = (double **)malloc(mr * sizeof(double *));
    for (int q = 0; q <= mr; q++) V[q] = dalloc(n);
    for (int q = 0; q < mr; q++) Z[q] = dalloc(n);
    double *H = dalloc((size_t)(mr + 1) * mr), *cs = dalloc(mr), *sn = dalloc(mr), *g = dalloc(mr + 1);
    double *w = dalloc(n), *y = dalloc(mr);
    for (size_t q = 0; q < n; q++) if (!L->row[q]) phi[q] = 0;
    double bnorm = sqrt(dot(L, b, b));
    if (bnorm == 0) bnorm = 1;
    int it = 0;
    double rel = 1;
    while (it < maxit) {
        residual(L, phi, b, V[0]);
        double beta = sqrt(dot(L, V[0], V[0]));
        rel = beta / bnorm;
        if (rel <= tol) break;
#pragma omp parallel for schedule(static)
        for (size_t q = 0; q < n; q++) V[0][q] /= beta;
        memset(g, 0, (mr + 1) * sizeof(double));
        g[0] = beta;
        int jj;
        for (jj = 0; jj < mr && it < maxit; jj++, it++) {
            /* z_j = M^-1 v_j (one V-cycle from zero) */
            memcpy(mg.b[0], V[jj], n * sizeof(double));
            memset(mg.x[0], 0, n * sizeof(double));
            if (levels > 1) vcycle(&mg, 0, 2, 2, 20);
            else for (int s = 0; s < 4; s++) smooth_lines(L, mg.x[0], mg.b[0]);
            memcpy(Z[jj], mg.x[0], n * sizeof(double));
            level_apply(L, Z[jj], w);
            /* modified Gram-Schmidt */
            for (int i = 0; i <= jj; i++) {
                double h = dot(L, w, V[i]);
                H[i * mr + jj] = h;
#pragma omp parallel for schedule(static)
                for (size_t q = 0; q < n; q++) w[q] -= h * V[i][q];
            }
            double hn = sqrt(dot(L, w, w));
            H[(jj + 1) * mr + jj] = hn;
#pragma omp parallel for schedule(static)
            for (size_t q = 0; q < n; q++) V[jj + 1][q] = hn > 0 ? w[q] / hn : 0;
            for (int i = 0; i < jj; i++) {
                double t = cs[i] * H[i * mr + jj] + sn[i] * H[(i + 1) * mr + jj];
                H[(i + 1) * mr + jj] = -sn[i] * H[i * mr + jj] + cs[i] * H[(i + 1) * mr + jj];
                H[i * mr + jj] = t;
            }
            double a = H[jj * mr + jj], bb = H[(jj + 1) * mr + jj];
            double den = sqrt(a * a + bb * bb);
            cs[jj] = a / den; sn[jj] = bb / den;
            H[jj * mr + jj] = den;
            H[(jj + 1) * mr + jj] = 0;
            g[jj + 1] = -sn[jj] * g[jj];
            g[jj] = cs[jj] * g[jj];
            if (fabs(g[jj + 1]) / bnorm <= tol) { jj++; it++; break; }
        }
        /* back substitution and update */
        for (int i = jj - 1; i >= 0; i--) {
            double s = g[i];
            for (int c = i + 1; c < jj; c++) s -= H[i * mr + c] * y[c];
            y[i] = s / H[i * mr + i];
        }
        for (int i = 0; i < jj; i++)
#pragma omp parallel for schedule(static)
            for (size_t q = 0; q < n; q++) phi[q] += y[i] * Z[i][q];
    }
    residual(L, phi, b, w);
    rel = sqrt(dot(L, w, w)) / bnorm;
    if (relres) *relres = rel;
    orc_remove_mean(p, m, phi);
    for (int q = 0; q <= mr; q++) free(V[q]);
    for (int q = 0; q < mr; q++) free(Z[q]);
    free(V); free(Z); free(H); free(cs); free(sn); free(g); free(w); free(y);
    mg_free(&mg);
    return it;
}

/* ------------------------------------------------------------------------ */
/* a13: Convert_Phi2_Phi periodic handling + UpdatePressure                  */
/* ------------------------------------------------------------------------ */

void orc_update_pressure(const orc_params *p, const orc_metrics *m, double *pres, double *phi)
{
    /* Convert_Phi2_Phi zeroes non-system cells and copies periodic ghosts */
    for (int k = 0; k < p->mz; k++)
        for (int j = 0; j < p->my; j++)
            for (int i = 0; i < p->mx; i++)
                if (!is_row(p, m, k, j, i)) S(phi, k, j, i) = 0;
    orc_periodic_fix(p, phi, 1);
    for (int k = 1; k < p->mz - 1; k++)
        for (int j = 1; j < p->my - 1; j++)
            for (int i = 1; i < p->mx - 1; i++) {
                if (S(m->nvert, k, j, i) < 0.1) S(pres, k, j, i) += S(phi, k, j, i);
                if (S(m->nvert, k, j, i) > 1.1) S(pres, k, j, i) = 0;
            }
    orc_periodic_fix(p, pres, 1);
}

/* ------------------------------------------------------------------------ */
/* a14: Projection (poisson.c:2700-3066)                                     */
/* ------------------------------------------------------------------------ */

void orc_projection(const orc_params *p, const orc_metrics *m, const double *phi, double *ucont,
                    double *ucat, const double *ucat_old)
{
    const int mx = p->mx, my = p->my, mz = p->mz;
    const double th = p->poisson_threshold;
    const double *nv = m->nvert;
    const int pi_ = p->per_i, pj = p->per_j, pk = p->per_k;
    const double coeff = 1.;
#define PH(kk, jj, ii) S(phi, kk, jj, ii)
#define NV(kk, jj, ii) S(nv, kk, jj, ii)
#pragma omp parallel for schedule(static)
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t q = IX(p, k, j, i);
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
                        const double *ic = &m->icsi[3 * q], *ie = &m->ieta[3 * q], *iz = &m->izet[3 * q];
                        double iaj = m->iaj[q];
                        VX(ucont, k, j, i) -= (dpdc * (ic[0] * ic[0] + ic[1] * ic[1] + ic[2] * ic[2]) * iaj +
                                               dpde * (ie[0] * ic[0] + ie[1] * ic[1] + ie[2] * ic[2]) * iaj +
                                               dpdz * (iz[0] * ic[0] + iz[1] * ic[1] + iz[2] * ic[2]) * iaj) *
                                              p->dt * p->st / coeff;
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
                        const double *jc = &m->jcsi[3 * q], *je = &m->jeta[3 * q], *jz = &m->jzet[3 * q];
                        double jaj = m->jaj[q];
                        VY(ucont, k, j, i) -= (dpdc * (jc[0] * je[0] + jc[1] * je[1] + jc[2] * je[2]) * jaj +
                                               dpde * (je[0] * je[0] + je[1] * je[1] + je[2] * je[2]) * jaj +
                                               dpdz * (jz[0] * je[0] + jz[1] * je[1] + jz[2] * je[2]) * jaj) *
                                              p->dt * p->st / coeff;
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
                        const double *kc = &m->kcsi[3 * q], *ke = &m->keta[3 * q], *kz = &m->kzet[3 * q];
                        double kaj = m->kaj[q];
                        VZ(ucont, k, j, i) -= (dpdc * (kc[0] * kz[0] + kc[1] * kz[1] + kc[2] * kz[2]) * kaj +
                                               dpde * (ke[0] * kz[0] + ke[1] * kz[1] + ke[2] * kz[2]) * kaj +
                                               dpdz * (kz[0] * kz[0] + kz[1] * kz[1] + kz[2] * kz[2]) * kaj) *
                                              p->dt * p->st / coeff;
                    }
                }
            }
#undef PH
#undef NV
    /* component-wise periodic fix, poisson.c:3001-3040 */
    if (pi_ || pj || pk) {
        for (int k = 0; k < mz; k++)
            for (int j = 0; j < my; j++)
                for (int i = 0; i < mx; i++) {
                    int a, b, c, fi, fj, fk;
                    if (!periodic_src(p, k, j, i, &c, &b, &a, &fi, &fj, &fk)) continue;
                    size_t d = IX(p, k, j, i), s = IX(p, c, b, a);
                    if (fi) ucont[3 * d + 0] = ucont[3 * s + 0];
                    if (fj) ucont[3 * d + 1] = ucont[3 * s + 1];
                    if (fk) ucont[3 * d + 2] = ucont[3 * s + 2];
                }
    }
    orc_contra2cart(p, m, ucont, ucat, ucat_old);
}

/* ------------------------------------------------------------------------ */
/* a15: Divergence (main.c:936-1074)                                         */
/* ------------------------------------------------------------------------ */

double orc_divergence(const orc_params *p, const orc_metrics *m, const double *ucont)
{
    double mx_ = 0;
    const double *nv = m->nvert;
    for (int k = 1; k < p->mz - 1; k++)
        for (int j = 1; j < p->my - 1; j++)
            for (int i = 1; i < p->mx - 1; i++) {
                double d = fabs((VX(ucont, k, j, i) - VX(ucont, k, j, i - 1) + VY(ucont, k, j, i) - VY(ucont, k, j - 1, i) +
                                 VZ(ucont, k, j, i) - VZ(ucont, k - 1, j, i)) * S(m->aj, k, j, i));
                if (S(nv, k, j, i) + S(nv, k + 1, j, i) + S(nv, k - 1, j, i) + S(nv, k, j + 1, i) + S(nv, k, j - 1, i) +
                        S(nv, k, j, i + 1) + S(nv, k, j, i - 1) > 0.1)
                    d = 0.;
                if (d > mx_) mx_ = d;
            }
    return mx_;
}

/* ------------------------------------------------------------------------ */
/* one fractional step in Flow_Solver order (solvers.c:313-759), -imp 0      */
/* ------------------------------------------------------------------------ */

int orc_fractional_step(orc_params *p, const orc_metrics *m, double *ucont, double *ucat,
                        const double *ucat_old, double *nu_t, double *pres, double *phi, double *dp,
                        double tol, double *maxdiv, double *relres)
{
    size_t n = npts(p);
    /* pseudo_periodic_BC, solvers.c:355-358: this step's inlet plane */
    const double *plane_in = p->inflow_plane;
    double *plane = NULL;
    if (p->pseudo_periodic) {
        plane = dalloc(3 * (size_t)p->mx * p->my);
        orc_pseudo_periodic_plane(p, ucat, plane);
        p->inflow_plane = plane;
    }
    /* Calc_k_Flux, solvers.c:363 */
    p->mean_k_flux = orc_calc_k_flux(p, m, ucont);
    /* LES, solvers.c:365-371 */
    if (p->les) {
        orc_contra2cart(p, m, ucont, ucat, ucat_old);
        if (p->les == 2) {
            /* dynamic_freq 1, after the first steps (ti >= 2) */
            double *cs = dalloc(n);
            orc_smagorinsky_constant(p, m, ucat, 2, 0, 0, cs);
            orc_eddy_viscosity_cs(p, m, ucat, cs, nu_t);
            free(cs);
        } else {
            orc_eddy_viscosity(p, m, ucat, nu_t);
        }
    }
    /* solvers.c:438 */
    orc_pressure_gradient(p, m, pres, dp);
    /* solvers.c:635 */
    orc_runge_kutta(p, m, ucont, ucat, ucat_old, nu_t, dp, NULL);
    /* solvers.c:650-655 */
    double *b = dalloc(n);
    orc_poisson_rhs(p, m, ucont, b);
    int it = orc_poisson_solve(p, m, b, phi, tol, 20, 400, 6, relres);
    free(b);
    /* solvers.c:660-666 */
    orc_update_pressure(p, m, pres, phi);
    orc_projection(p, m, phi, ucont, ucat, ucat_old);
    /* solvers.c:684 */
    if (maxdiv) *maxdiv = orc_divergence(p, m, ucont);
    if (plane) { p->inflow_plane = plane_in; free(plane); }
    return it;
}
