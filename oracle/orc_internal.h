/*
 * orc_internal.h -- indexing helpers shared by the oracle's translation units.
 * TEST INFRASTRUCTURE ONLY (see vfs_oracle.h).
 */
#ifndef ORC_INTERNAL_H
#define ORC_INTERNAL_H

#include "vfs_oracle.h"

#include <assert.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* indexing                                                                  */
/* ------------------------------------------------------------------------ */

static inline int wrapi(int i, int m, int per)
{
    if (per) {
        if (i < 0) i += m;
        else if (i >= m) i -= m;
    }
    return i;
}

static inline size_t IX(const orc_params *p, int k, int j, int i)
{
    i = wrapi(i, p->mx, p->per_i);
    j = wrapi(j, p->my, p->per_j);
    k = wrapi(k, p->mz, p->per_k);
    assert(i >= 0 && i < p->mx && j >= 0 && j < p->my && k >= 0 && k < p->mz);
    return ((size_t)k * p->my + j) * p->mx + i;
}

#define S(a, k, j, i) ((a)[IX(p, (k), (j), (i))])
#define VX(a, k, j, i) ((a)[3 * IX(p, (k), (j), (i)) + 0])
#define VY(a, k, j, i) ((a)[3 * IX(p, (k), (j), (i)) + 1])
#define VZ(a, k, j, i) ((a)[3 * IX(p, (k), (j), (i)) + 2])
#define VC(a, c, k, j, i) ((a)[3 * IX(p, (k), (j), (i)) + (c)])

/* Calculate_Covariant_metrics (rhs2.c:595-612): G = inverse of [csi; eta; zet]. */
void orc_covariant(const double g[3][3], double G[3][3]);

static inline size_t npts(const orc_params *p) { return (size_t)p->mx * p->my * p->mz; }

static inline double *dalloc(size_t n)
{
    double *a = (double *)calloc(n ? n : 1, sizeof(double));
    if (!a) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return a;
}

/* Source cell of the periodic fix for (k,j,i); returns 0 when the cell is not
 * a periodic ghost (e.g. momentum.c:638-666, rhs.c:129-156). */
static inline int periodic_src(const orc_params *p, int k, int j, int i, int *c, int *b, int *a,
                        int *fi, int *fj, int *fk)
{
    int flag = 0;
    *a = i; *b = j; *c = k;
    *fi = *fj = *fk = 0;
    if (p->per_i && i == 0) { *a = -2; *fi = 1; flag = 1; }
    else if (p->per_i && i == p->mx - 1) { *a = p->mx + 1; *fi = 1; flag = 1; }
    if (p->per_j && j == 0) { *b = -2; *fj = 1; flag = 1; }
    else if (p->per_j && j == p->my - 1) { *b = p->my + 1; *fj = 1; flag = 1; }
    if (p->per_k && k == 0) { *c = -2; *fk = 1; flag = 1; }
    else if (p->per_k && k == p->mz - 1) { *c = p->mz + 1; *fk = 1; flag = 1; }
    return flag;
}

#endif
