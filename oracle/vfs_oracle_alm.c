/*
 * vfs_oracle_alm.c -- actuator-line delta kernels (rotor_model.c:5095-5393).
 * TEST INFRASTRUCTURE ONLY (see vfs_oracle.h).
 *
 * The truncated pi constants are the reference's own (3.14159265 in the
 * s3h/s4h kernels, 3.1415926 in dfunc_exp): matching them keeps parity at
 * round-off level (the difference to M_PI is ~1e-9 relative).
 */
#include "orc_internal.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846 /* math.h M_PI, used by refAL_Rot */
#endif

/* dfunc_2h, rotor_model.c:5130-5137 */
double orc_dfunc_2h(double r)
{
    if (fabs(r) < 1.0) return 1.0 - fabs(r);
    return 0.0;
}

/* dfunc_s3h, rotor_model.c:5117-5127 */
double orc_dfunc_s3h(double r)
{
    double a = fabs(r);
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

/* dfunc_s4h, rotor_model.c:5140-5154 */
double orc_dfunc_s4h(double r)
{
    double a = fabs(r);
    if (a <= 0.5) return 3.0 / 8.0 + 3.14159265 / 32.0 - pow(r, 2) / 4.0;
    if (a >= 0.5 && a <= 1.5)
        return 1.0 / 4.0 + (1.0 - a) * sqrt(-2.0 + 8.0 * a - 4.0 * pow(r, 2)) / 8.0 -
               asin(sqrt(2.0) * (a - 1.0)) / 8.0;
    if (a >= 1.5 && a <= 2.5)
        return 17.0 / 16.0 - 3.14159265 / 64.0 - 3.0 * a / 4.0 + pow(r, 2) / 8.0 +
               (a - 2.0) * sqrt(-14.0 + 16.0 * a - 4.0 * pow(r, 2)) / 16.0 +
               asin(sqrt(2.0) * (a - 2.0)) / 16.0;
    return 0.0;
}

/* dfunc_exp, rotor_model.c:5230-5234 */
double orc_dfunc_exp(double r, double n)
{
    return exp(-(r / n) * (r / n)) / (pow(n, 1) * pow(3.1415926, 0.5));
}

/* ------------------------------------------------------------------------ */
/* actuator line (rotor_model.c, rotor_model 3/6 with rotor_Rot)             */
/* ------------------------------------------------------------------------ */

/* ArbitraryRotate, rotor_model.c:5174-5192: p rotated by theta about r. */
void orc_arbitrary_rotate(const double p[3], double theta, const double r_in[3], double q[3])
{
    double r[3] = {r_in[0], r_in[1], r_in[2]};
    double rr = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) + 1.0e-11;
    r[0] = r[0] / rr; r[1] = r[1] / rr; r[2] = r[2] / rr;
    double c = cos(theta), s = sin(theta);
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

/* Line elements of nb blades with nvb nodes each (element i = blade b, node
 * pair (b*nvb+j, b*nvb+j+1)): centre and length, rotor_model.c:4838-4864. */
void orc_alm_elements(int nb, int nvb, const double *xbp, double *cent, double *dA)
{
    int ne = nvb - 1;
    for (int b = 0; b < nb; b++)
        for (int j = 0; j < ne; j++) {
            int i = b * ne + j, n1 = b * nvb + j, n2 = n1 + 1;
            double dx = xbp[3 * n2] - xbp[3 * n1], dy = xbp[3 * n2 + 1] - xbp[3 * n1 + 1], dz = xbp[3 * n2 + 2] - xbp[3 * n1 + 2];
            dA[i] = sqrt(dx * dx + dy * dy + dz * dz);
            cent[3 * i] = (xbp[3 * n1] + xbp[3 * n2]) / 2.;
            cent[3 * i + 1] = (xbp[3 * n1 + 1] + xbp[3 * n2 + 1]) / 2.;
            cent[3 * i + 2] = (xbp[3 * n1 + 2] + xbp[3 * n2 + 2]) / 2.;
        }
}

/* rotor_Rot node loop, rotor_model.c:4775-4806: rotate every node by
 * angvel*dt about the axis through the hub and set its velocity
 * angvel*|r| * (axis x r/|r|). */
void orc_alm_rotate(const orc_turbine *t, double dt, int nv, double *xbp, double *unode)
{
    const double na[3] = {t->nx, t->ny, t->nz};
    for (int i = 0; i < nv; i++) {
        double p[3] = {xbp[3 * i] - t->xc, xbp[3 * i + 1] - t->yc, xbp[3 * i + 2] - t->zc}, q[3];
        orc_arbitrary_rotate(p, t->angvel * dt, na, q);
        xbp[3 * i] = q[0] + t->xc;
        xbp[3 * i + 1] = q[1] + t->yc;
        xbp[3 * i + 2] = q[2] + t->zc;
        double rx = xbp[3 * i] - t->xc, ry = xbp[3 * i + 1] - t->yc, rz = xbp[3 * i + 2] - t->zc;
        double rr = sqrt(rx * rx + ry * ry + rz * rz) + 1.e-19;
        double nr[3] = {rx / rr, ry / rr, rz / rr};
        double nt[3] = {na[1] * nr[2] - na[2] * nr[1], na[2] * nr[0] - na[0] * nr[2], na[0] * nr[1] - na[1] * nr[0]};
        double ut = t->angvel * rr;
        unode[3 * i] = ut * nt[0];
        unode[3 * i + 1] = ut * nt[1];
        unode[3 * i + 2] = ut * nt[2];
    }
}

/* refAL_Rot node loop, rotor_model.c:5046-5062: the reference line lies
 * refangle degrees ahead in the direction of rotation. */
void orc_alm_refline(const orc_turbine *t, int nv, const double *xbp, double *xref)
{
    const double na[3] = {t->nx, t->ny, t->nz};
    double isign = (1.e-19 + t->angvel) / (fabs(t->angvel) + 1.e-19);
    isign = isign > 0.0 ? 1.0 : -1.0;
    double theta = isign * t->refangle * M_PI / 180.0;
    for (int i = 0; i < nv; i++) {
        double p[3] = {xbp[3 * i] - t->xc, xbp[3 * i + 1] - t->yc, xbp[3 * i + 2] - t->zc}, q[3];
        orc_arbitrary_rotate(p, theta, na, q);
        xref[3 * i] = q[0] + t->xc;
        xref[3 * i + 1] = q[1] + t->yc;
        xref[3 * i + 2] = q[2] + t->zc;
    }
}

/* Twist and chord by radius, airfoil_ACL rotor_model.c:2118-2134. */
void orc_alm_twist_chord(const orc_turbine *t, int n, const double *cent, const orc_foil *f, int nfoil,
                         double *twist, double *chord)
{
    for (int i = 0; i < n; i++) {
        double r = sqrt(pow(cent[3 * i] - t->xc, 2.0) + pow(cent[3 * i + 1] - t->yc, 2.0) + pow(cent[3 * i + 2] - t->zc, 2.0));
        for (int q = 0; q < nfoil; q++) {
            if (!(r >= f[q].r_beg && r <= f[q].r_end)) continue;
            for (int j = 0; j < f[q].n_ac - 1; j++)
                if (r >= f[q].r_ac[j] && r <= f[q].r_ac[j + 1]) {
                    double fac1 = (f[q].r_ac[j + 1] - r) / (f[q].r_ac[j + 1] - f[q].r_ac[j]);
                    double fac2 = (-f[q].r_ac[j] + r) / (f[q].r_ac[j + 1] - f[q].r_ac[j]);
                    twist[i] = fac1 * f[q].twist[j] + fac2 * f[q].twist[j + 1];
                    chord[i] = fac1 * f[q].chord[j] + fac2 * f[q].chord[j + 1];
                }
        }
    }
}

/* cell spacings along the three grid directions, rotor_model.c:1935-1940 */
static void cell_spacing(const orc_params *p, const orc_metrics *m, int k, int j, int i, double *hx, double *hy, double *hz)
{
    double a = sqrt(VX(m->csi, k, j, i) * VX(m->csi, k, j, i) + VY(m->csi, k, j, i) * VY(m->csi, k, j, i) +
                    VZ(m->csi, k, j, i) * VZ(m->csi, k, j, i));
    *hx = 1.0 / S(m->aj, k, j, i) / a;
    a = sqrt(VX(m->eta, k, j, i) * VX(m->eta, k, j, i) + VY(m->eta, k, j, i) * VY(m->eta, k, j, i) +
             VZ(m->eta, k, j, i) * VZ(m->eta, k, j, i));
    *hy = 1.0 / S(m->aj, k, j, i) / a;
    a = sqrt(VX(m->zet, k, j, i) * VX(m->zet, k, j, i) + VY(m->zet, k, j, i) * VY(m->zet, k, j, i) +
             VZ(m->zet, k, j, i) * VZ(m->zet, k, j, i));
    *hz = 1.0 / S(m->aj, k, j, i) / a;
}

/* Calculate_normal, rhs2.c:614-647: unit covariant base vectors */
static void cell_normals(const orc_params *p, const orc_metrics *m, int k, int j, int i, double ni[3], double nj[3],
                         double nk[3])
{
    double g[3][3], G[3][3];
    for (int c = 0; c < 3; c++) {
        g[0][c] = VC(m->csi, c, k, j, i);
        g[1][c] = VC(m->eta, c, k, j, i);
        g[2][c] = VC(m->zet, c, k, j, i);
    }
    orc_covariant(g, G);
    double x[3][3];
    for (int d = 0; d < 3; d++)
        for (int c = 0; c < 3; c++) x[d][c] = G[c][d];
    for (int d = 0; d < 3; d++) {
        double s = sqrt(x[d][0] * x[d][0] + x[d][1] * x[d][1] + x[d][2] * x[d][2]);
        x[d][0] /= s; x[d][1] /= s; x[d][2] /= s;
    }
    for (int c = 0; c < 3; c++) { ni[c] = x[0][c]; nj[c] = x[1][c]; nk[c] = x[2][c]; }
}

#define CO(c, k, j, i) (m->coor[3 * IX(p, (k), (j), (i)) + (c)])

/* Pre_process, rotor_model.c:1849-2011 (one rank): nearest grid node of each
 * element centre along the first interior grid lines, kept when closer than
 * 2*|(dhx,dhy,dhz)| of that cell, and the box [c-6, c+6) clipped to the
 * interior.  box = imin, imax, jmin, jmax, kmin, kmax per element.  An
 * element that is not close to any node gets index 0 (the single-rank
 * reference divides by a zero count there). */
void orc_alm_preprocess(const orc_params *p, const orc_metrics *m, int n, const double *cent, int *box)
{
    const int W = 6; /* Itpwidth = (int)halfwidth_dfunc + 2 with the default 4.0, rotor_model.c:47 */
    const int lxs = 1, lxe = p->mx - 1, lys = 1, lye = p->my - 1, lzs = 1, lze = p->mz - 1;
    for (int l = 0; l < n; l++) {
        const double cx = cent[3 * l], cy = cent[3 * l + 1], cz = cent[3 * l + 2];
        double dmin = 1.e6;
        int ic = 0, jc = 0, kc = 0;
        for (int i = lxs; i < lxe; i++) {
            int j = lys + 1, k = lzs + 1;
            double r1 = cx - CO(0, k, j, i), r2 = cy - CO(1, k, j, i), r3 = cz - CO(2, k, j, i);
            double d1 = sqrt(r1 * r1 + r2 * r2 + r3 * r3);
            if (d1 < dmin) { dmin = d1; ic = i; }
        }
        dmin = 1.e6;
        for (int j = lys; j < lye; j++) {
            int k = lzs + 1, i = lxs + 1;
            double r1 = cx - CO(0, k, j, i), r2 = cy - CO(1, k, j, i), r3 = cz - CO(2, k, j, i);
            double d1 = sqrt(r1 * r1 + r2 * r2 + r3 * r3);
            if (d1 < dmin) { dmin = d1; jc = j; }
        }
        dmin = 1.e6;
        for (int k = lzs; k < lze; k++) {
            int i = lxs + 1, j = lys + 1;
            double r1 = cx - CO(0, k, j, i), r2 = cy - CO(1, k, j, i), r3 = cz - CO(2, k, j, i);
            double d1 = sqrt(r1 * r1 + r2 * r2 + r3 * r3);
            if (d1 < dmin) { dmin = d1; kc = k; }
        }
        double r1 = cx - CO(0, kc, jc, ic), r2 = cy - CO(1, kc, jc, ic), r3 = cz - CO(2, kc, jc, ic);
        dmin = sqrt(r1 * r1 + r2 * r2 + r3 * r3);
        double hx, hy, hz;
        cell_spacing(p, m, kc, jc, ic, &hx, &hy, &hz);
        double dh_grid = 2.0 * sqrt(hx * hx + hy * hy + hz * hz);
        if (dmin - dh_grid > 1.e-6) ic = jc = kc = 0;
        int c[3] = {ic, jc, kc}, lo[3] = {lxs, lys, lzs}, hi[3] = {lxe, lye, lze};
        for (int d = 0; d < 3; d++) {
            int c1 = c[d] - W, c2 = c[d] + W;
            if (c1 > hi[d] || c2 < lo[d]) { box[6 * l + 2 * d] = lo[d]; box[6 * l + 2 * d + 1] = lo[d]; }
            else {
                box[6 * l + 2 * d] = c1 > lo[d] ? c1 : lo[d];
                box[6 * l + 2 * d + 1] = c2 < hi[d] ? c2 : hi[d];
            }
        }
    }
}

/* Calc_U_lagr, rotor_model.c:2937-3036 (one rank, no periodic turbines,
 * no moving frame): delta-weighted sum of ucat over the element's box with
 * dfunc_s3h of the cell-centre offset along the unit normals over the
 * local spacing. */
void orc_alm_ulagr(const orc_params *p, const orc_metrics *m, const double *ucat, int n, const double *cent,
                   const int *box, double *ulagr)
{
    for (int l = 0; l < n; l++) {
        double u[3] = {0, 0, 0};
        const int *b = box + 6 * l;
        for (int k = b[4]; k < b[5]; k++)
            for (int j = b[2]; j < b[3]; j++)
                for (int i = b[0]; i < b[1]; i++) {
                    double xcc[3];
                    for (int c = 0; c < 3; c++)
                        xcc[c] = 0.125 * (CO(c, k, j, i) + CO(c, k, j - 1, i) + CO(c, k - 1, j, i) + CO(c, k - 1, j - 1, i) +
                                          CO(c, k, j, i - 1) + CO(c, k, j - 1, i - 1) + CO(c, k - 1, j, i - 1) +
                                          CO(c, k - 1, j - 1, i - 1));
                    double hx, hy, hz, ni[3], nj[3], nk[3];
                    cell_spacing(p, m, k, j, i, &hx, &hy, &hz);
                    double rx = xcc[0] - cent[3 * l], ry = xcc[1] - cent[3 * l + 1], rz = xcc[2] - cent[3 * l + 2];
                    cell_normals(p, m, k, j, i, ni, nj, nk);
                    double r1 = (rx * ni[0] + ry * ni[1] + rz * ni[2]) / hx;
                    double r2 = (rx * nj[0] + ry * nj[1] + rz * nj[2]) / hy;
                    double r3 = (rx * nk[0] + ry * nk[1] + rz * nk[2]) / hz;
                    double df = orc_dfunc_s3h(r1) * orc_dfunc_s3h(r2) * orc_dfunc_s3h(r3);
                    u[0] += VX(ucat, k, j, i) * df;
                    u[1] += VY(ucat, k, j, i) * df;
                    u[2] += VZ(ucat, k, j, i) * df;
                }
        ulagr[3 * l] = u[0];
        ulagr[3 * l + 1] = u[1];
        ulagr[3 * l + 2] = u[2];
    }
}

/* polar lookup with the reference's out-of-range rules, rotor_model.c:2426-2461 */
static double polar(double aoa, int nq, const double *ang, const double *val, double lim, int lift, double oldv)
{
    const double pi = 3.141592653589793;
    double v = oldv;
    int inrange = 0;
    for (int j = 0; j < nq - 1; j++)
        if (aoa >= ang[j] && aoa <= ang[j + 1]) {
            double fac1 = (ang[j + 1] - aoa) / (ang[j + 1] - ang[j]);
            double fac2 = (-ang[j] + aoa) / (ang[j + 1] - ang[j]);
            v = fac1 * val[j] + fac2 * val[j + 1];
            inrange = 1;
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

/* Calc_F_lagr_ACL, rotor_model.c:2370-2597 for rotor_model 3/6 without the
 * optional tip/3D corrections and force smoothing (all off by default):
 * relative velocity minus its radial part, angle of attack, polar lookup,
 * drag along the relative velocity and lift along n_L, scaled by the chord;
 * elements inside the nacelle radius carry no force.  cl/cd hold the
 * previous values on entry (used when no foil covers the radius). */
void orc_alm_flagr(const orc_turbine *t, int nb, int nvb, const double *cent, const double *ulagr, const double *unode,
                   const double *twist, const double *chord, const orc_foil *f, int nfoil, double *aoa, double *cl,
                   double *cd, double *flagr)
{
    const double pi = 3.141592653589793;
    const int ne = nvb - 1;
    for (int b = 0; b < nb; b++)
        for (int jj = 0; jj < ne; jj++) {
            const int l = b * ne + jj, n1 = b * nvb + jj, n2 = n1 + 1;
            double rx = cent[3 * l] - t->xc, ry = cent[3 * l + 1] - t->yc, rz = cent[3 * l + 2] - t->zc;
            double r = sqrt(rx * rx + ry * ry + rz * rz) + 1.e-20;
            double nbl[3] = {rx / r, ry / r, rz / r};
            double ux = 0.5 * (unode[3 * n1] + unode[3 * n2]);
            double uy = 0.5 * (unode[3 * n1 + 1] + unode[3 * n2 + 1]);
            double uz = 0.5 * (unode[3 * n1 + 2] + unode[3 * n2 + 2]);
            double ublade = sqrt(pow(ux, 2) + pow(uy, 2) + pow(uz, 2));
            double nrot[3] = {ux / (ublade + 1.e-20), uy / (ublade + 1.e-20), uz / (ublade + 1.e-20)};
            double ub[3] = {ulagr[3 * l] - ux, ulagr[3 * l + 1] - uy, ulagr[3 * l + 2] - uz};
            double ur = ub[0] * nbl[0] + ub[1] * nbl[1] + ub[2] * nbl[2];
            ub[0] -= ur * nbl[0];
            ub[1] -= ur * nbl[1];
            ub[2] -= ur * nbl[2];
            double uref = sqrt(pow(ub[0], 2) + pow(ub[1], 2) + pow(ub[2], 2));
            double nrel[3] = {ub[0] / (uref + 1.e-20), ub[1] / (uref + 1.e-20), ub[2] / (uref + 1.e-20)};
            double tmp = -nrel[0] * nrot[0] - nrel[1] * nrot[1] - nrel[2] * nrot[2];
            tmp /= (1 + 1e-9);
            aoa[l] = acos(tmp) * 180.0 / pi - twist[l] - t->pitch;
            double nL[3] = {nrel[1] * nbl[2] - nrel[2] * nbl[1], nrel[2] * nbl[0] - nrel[0] * nbl[2],
                            nrel[0] * nbl[1] - nrel[1] * nbl[0]};
            tmp = t->nx * nL[0] + t->ny * nL[1] + t->nz * nL[2];
            if (tmp < 0.0) { nL[0] = -nL[0]; nL[1] = -nL[1]; nL[2] = -nL[2]; }
            for (int q = 0; q < nfoil; q++)
                if (r >= f[q].r_beg && r <= f[q].r_end) {
                    cd[l] = polar(aoa[l], f[q].n_cd, f[q].ang_cd, f[q].cd, 1.2, 0, cd[l]);
                    cl[l] = polar(aoa[l], f[q].n_cl, f[q].ang_cl, f[q].cl, 1.05, 1, cl[l]);
                }
            const double factor = chord[l];
            for (int c = 0; c < 3; c++) flagr[3 * l + c] = -0.5 * uref * uref * fabs(cd[l]) * nrel[c] * factor;
            for (int c = 0; c < 3; c++) flagr[3 * l + c] += -0.5 * uref * uref * fabs(cl[l]) * nL[c] * factor;
            /* tip-loss factor F1, rotor_model.c:2488-2544 (the later model wins; its pi is 3.1415926) */
            double F1 = 1.0;
            if (t->shen) {
                const double R = t->r_rotor, pi1 = 3.1415926;
                const double phi1 = (aoa[l] + twist[l] + t->pitch) * pi1 / 180.0;
                const double g1 = exp(-t->a_shen * (t->num_blade * fabs(t->tsr) - t->b_shen)) + t->c_shen;
                F1 = 2.0 * acos(exp(-g1 * t->num_blade * (fmax(R - r, 0.0)) / 2.0 / r / fabs(sin(phi1)))) / pi1;
            }
            if (t->prandtl) {
                const double R = t->r_rotor, pi1 = 3.1415926;
                const double g1 = sqrt(1 + pow(t->tsr, 2));
                F1 = 2.0 * acos(exp(-g1 * t->num_blade * (fmax(R - r, 0.0)) / 2.0 / R)) / pi1;
            }
            if (t->shen1) {
                const double R = t->r_rotor, pi1 = 3.1415926;
                const double phi1 = (aoa[l] + twist[l] + t->pitch) * pi1 / 180.0;
                const double Lambda = fabs(t->tsr);
                double g1;
                if (Lambda < 7.0636) g1 = -1.7636 * Lambda + 11.6112;
                else if (Lambda > 7.0636 && Lambda < 9.8146) g1 = -0.5805 * Lambda + 3.2542;
                else g1 = -0.5077 * Lambda + 2.5397;
                g1 *= t->correction_shen1;
                g1 = exp(g1);
                F1 = 2.0 * acos(exp(-g1 * t->num_blade * (fmax(R - r, 0.0)) / 2.0 / r / fabs(sin(phi1)))) / pi1;
            }
            for (int c = 0; c < 3; c++) flagr[3 * l + c] *= F1;
            /* 3D lift correction inboard of 0.75 R along the rotor axis, rotor_model.c:2548-2558 */
            if (t->correction3d_cl && r < 0.75 * t->r_rotor) {
                const double inflow = twist[l] * pi / 180 + aoa[l] * pi / 180;
                const double Fc = -0.5 * t->c0_cl * pow(chord[l] / r, 1) * pow(cos(inflow), 2) * uref * uref * factor;
                flagr[3 * l] += Fc * t->nx;
                flagr[3 * l + 1] += Fc * t->ny;
                flagr[3 * l + 2] += Fc * t->nz;
            }
        }
    /* force smoothing along each blade, rotor_model.c:2562-2598: smoothforce + 1
     * passes (do-while), each element j in [4, ne - 4) of a blade takes
     * sum_{i = j-4}^{j+3} F_i dfunc_s4h(i - j) of the previous pass */
    if (t->smoothforce) {
        const int n = nb * ne;
        double *tmp = (double *)malloc(3 * (size_t)n * sizeof(double));
        int pass = 0;
        do {
            pass++;
            memcpy(tmp, flagr, 3 * (size_t)n * sizeof(double));
            for (int b = 0; b < nb; b++)
                for (int j = b * ne + 4; j < (b + 1) * ne - 4; j++) {
                    double s[3] = {0, 0, 0};
                    for (int i = j - 4; i < j + 4; i++) {
                        const double w = orc_dfunc_s4h((double)(i - j));
                        for (int c = 0; c < 3; c++) s[c] += tmp[3 * i + c] * w;
                    }
                    for (int c = 0; c < 3; c++) flagr[3 * j + c] = s[c];
                }
        } while (pass <= t->smoothforce);
        free(tmp);
    }
    for (int l = 0; l < nb * ne; l++) {
        double r = sqrt(pow((cent[3 * l] - t->xc), 2) + pow((cent[3 * l + 1] - t->yc), 2) + pow((cent[3 * l + 2] - t->zc), 2));
        if (r < t->r_nacelle) flagr[3 * l] = flagr[3 * l + 1] = flagr[3 * l + 2] = 0.0;
    }
}

/* Calc_F_eul, rotor_model.c:3668-3929 (forcewidthfixed 0): every element adds
 * F_lagr * delta * dA projected on csi/eta/zet to the cells of its box, the
 * delta evaluated at the cell's i/j/k face centres (df 0: dfunc_2h, 7:
 * dfunc_exp with halfwidth, otherwise dfunc_s4h) times 1/(dhx dhy dhz);
 * then cells next to blanking (sum of 27 nvert > 2.9) and the first and
 * last interior layers with their ghosts are zeroed.  feul accumulates. */
void orc_alm_feul(const orc_params *p, const orc_metrics *m, int n, const double *cent, const int *box,
                  const double *flagr, const double *dA, int df, double halfwidth, double *feul)
{
    for (int l = 0; l < n; l++) {
        const int *b = box + 6 * l;
        for (int k = b[4]; k < b[5]; k++)
            for (int j = b[2]; j < b[3]; j++)
                for (int i = b[0]; i < b[1]; i++) {
                    double fi[3], fj[3], fk[3];
                    for (int c = 0; c < 3; c++) {
                        fi[c] = (CO(c, k, j, i) + CO(c, k - 1, j, i) + CO(c, k, j - 1, i) + CO(c, k - 1, j - 1, i)) * 0.25;
                        fj[c] = (CO(c, k, j, i) + CO(c, k - 1, j, i) + CO(c, k, j, i - 1) + CO(c, k - 1, j, i - 1)) * 0.25;
                        fk[c] = (CO(c, k, j, i) + CO(c, k, j - 1, i) + CO(c, k, j, i - 1) + CO(c, k, j - 1, i - 1)) * 0.25;
                    }
                    double ri[3] = {fi[0] - cent[3 * l], fi[1] - cent[3 * l + 1], fi[2] - cent[3 * l + 2]};
                    double rj[3] = {fj[0] - cent[3 * l], fj[1] - cent[3 * l + 1], fj[2] - cent[3 * l + 2]};
                    double rk[3] = {fk[0] - cent[3 * l], fk[1] - cent[3 * l + 1], fk[2] - cent[3 * l + 2]};
                    double ni[3], nj[3], nk[3], hx, hy, hz;
                    cell_normals(p, m, k, j, i, ni, nj, nk);
                    cell_spacing(p, m, k, j, i, &hx, &hy, &hz);
                    double r1i = fabs(ri[0] * ni[0] + ri[1] * ni[1] + ri[2] * ni[2]) / hx;
                    double r2i = fabs(ri[0] * nj[0] + ri[1] * nj[1] + ri[2] * nj[2]) / hy;
                    double r3i = fabs(ri[0] * nk[0] + ri[1] * nk[1] + ri[2] * nk[2]) / hz;
                    double r1j = fabs(rj[0] * ni[0] + rj[1] * ni[1] + rj[2] * ni[2]) / hx;
                    double r2j = fabs(rj[0] * nj[0] + rj[1] * nj[1] + rj[2] * nj[2]) / hy;
                    double r3j = fabs(rj[0] * nk[0] + rj[1] * nk[1] + rj[2] * nk[2]) / hz;
                    double r1k = fabs(rk[0] * ni[0] + rk[1] * ni[1] + rk[2] * ni[2]) / hx;
                    double r2k = fabs(rk[0] * nj[0] + rk[1] * nj[1] + rk[2] * nj[2]) / hy;
                    double r3k = fabs(rk[0] * nk[0] + rk[1] * nk[1] + rk[2] * nk[2]) / hz;
                    double vol = 1.0 / (hx * hy * hz);
                    double di, dj, dk;
                    if (df == 0) {
                        di = vol * orc_dfunc_2h(r1i) * orc_dfunc_2h(r2i) * orc_dfunc_2h(r3i);
                        dj = vol * orc_dfunc_2h(r1j) * orc_dfunc_2h(r2j) * orc_dfunc_2h(r3j);
                        dk = vol * orc_dfunc_2h(r1k) * orc_dfunc_2h(r2k) * orc_dfunc_2h(r3k);
                    } else if (df == 7) {
                        double hw = halfwidth;
                        di = vol * orc_dfunc_exp(r1i, hw) * orc_dfunc_exp(r2i, hw) * orc_dfunc_exp(r3i, hw);
                        dj = vol * orc_dfunc_exp(r1j, hw) * orc_dfunc_exp(r2j, hw) * orc_dfunc_exp(r3j, hw);
                        dk = vol * orc_dfunc_exp(r1k, hw) * orc_dfunc_exp(r2k, hw) * orc_dfunc_exp(r3k, hw);
                    } else {
                        di = vol * orc_dfunc_s4h(r1i) * orc_dfunc_s4h(r2i) * orc_dfunc_s4h(r3i);
                        dj = vol * orc_dfunc_s4h(r1j) * orc_dfunc_s4h(r2j) * orc_dfunc_s4h(r3j);
                        dk = vol * orc_dfunc_s4h(r1k) * orc_dfunc_s4h(r2k) * orc_dfunc_s4h(r3k);
                    }
                    const double Fx = flagr[3 * l], Fy = flagr[3 * l + 1], Fz = flagr[3 * l + 2], a = dA[l];
                    VX(feul, k, j, i) += Fx * di * a * VX(m->csi, k, j, i) + Fy * di * a * VY(m->csi, k, j, i) +
                                         Fz * di * a * VZ(m->csi, k, j, i);
                    VY(feul, k, j, i) += Fx * dj * a * VX(m->eta, k, j, i) + Fy * dj * a * VY(m->eta, k, j, i) +
                                         Fz * dj * a * VZ(m->eta, k, j, i);
                    VZ(feul, k, j, i) += Fx * dk * a * VX(m->zet, k, j, i) + Fy * dk * a * VY(m->zet, k, j, i) +
                                         Fz * dk * a * VZ(m->zet, k, j, i);
                }
    }
    for (int k = 1; k < p->mz - 1; k++)
        for (int j = 1; j < p->my - 1; j++)
            for (int i = 1; i < p->mx - 1; i++) {
                double nv = 0.0;
                for (int kk = k - 1; kk < k + 2; kk++)
                    for (int jj = j - 1; jj < j + 2; jj++)
                        for (int ii = i - 1; ii < i + 2; ii++) nv += S(m->nvert, kk, jj, ii);
                if (nv > 2.9) VX(feul, k, j, i) = VY(feul, k, j, i) = VZ(feul, k, j, i) = 0.0;
                int zs[6][3] = {{k, j, i - 1}, {k, j - 1, i}, {k - 1, j, i}, {k, j, i + 1}, {k, j + 1, i}, {k + 1, j, i}};
                int hit[6] = {i == 1, j == 1, k == 1, i == p->mx - 2, j == p->my - 2, k == p->mz - 2};
                for (int q = 0; q < 6; q++)
                    if (hit[q]) {
                        size_t a = IX(p, zs[q][0], zs[q][1], zs[q][2]), c = IX(p, k, j, i);
                        feul[3 * a] = feul[3 * a + 1] = feul[3 * a + 2] = 0.0;
                        feul[3 * c] = feul[3 * c + 1] = feul[3 * c + 2] = 0.0;
                    }
            }
}
#undef CO

/* Calc_forces_ACL, rotor_model.c:4148-4180: rotor sums over the elements */
void orc_alm_loads(const orc_turbine *t, int n, const double *cent, const double *flagr, const double *ulagr,
                   const double *dA, double out[8])
{
    double A_Sum = 0, F_Sum = 0, P_Sum = 0, U_Sum = 0, Fx = 0, Fy = 0, Fz = 0, M_Sum = 0;
    const double nx = t->nx, ny = t->ny, nz = t->nz;
    for (int l = 0; l < n; l++) {
        const double *F = flagr + 3 * l, *U = ulagr + 3 * l;
        A_Sum += dA[l];
        const double F_axis = F[0] * nx + F[1] * ny + F[2] * nz;
        const double U_axis = U[0] * nx + U[1] * ny + U[2] * nz;
        F_Sum += F_axis * dA[l];
        P_Sum += F_axis * dA[l] * U_axis;
        U_Sum += U_axis * dA[l];
        Fx += F[0] * dA[l];
        Fy += F[1] * dA[l];
        Fz += F[2] * dA[l];
        const double rx = cent[3 * l] - t->xc, ry = cent[3 * l + 1] - t->yc, rz = cent[3 * l + 2] - t->zc;
        const double M_x = ry * F[2] * dA[l] - rz * F[1] * dA[l];
        const double M_y = rz * F[0] * dA[l] - rx * F[2] * dA[l];
        const double M_z = rx * F[1] * dA[l] - ry * F[0] * dA[l];
        M_Sum += M_x * nx + M_y * ny + M_z * nz;
    }
    out[0] = A_Sum; out[1] = F_Sum; out[2] = P_Sum; out[3] = U_Sum / A_Sum;
    out[4] = Fx; out[5] = Fy; out[6] = Fz; out[7] = M_Sum;
}

/* Calc_turbineangvel, rotor_model.c:5519-5601 */
void orc_turbine_angvel(int mode, const double prm[13], double dt, double uref, int ti, int tistart, double st[5])
{
    const double angvel_fixed = prm[0], tsr_in = prm[1], r_rotor = prm[2], reflength_wt = prm[3];
    const double refvel_wt = prm[4], refvel_cfd = prm[5], rho_air = prm[6], J = prm[7], CP_max = prm[8];
    const double TSR_max = prm[9], X_control = prm[10], Tgen_fixed = prm[12];
    const int fixgen = prm[11] != 0;
    double R0 = r_rotor / reflength_wt;
    double rho_cfd = 1.0;
    double T_wt = r_rotor / refvel_wt;
    double T_cfd = R0 / refvel_cfd;
    double ratio_rho = rho_air / rho_cfd;
    double ratio_L = reflength_wt;
    double ratio_T = T_wt / T_cfd;
    double ratio_V = refvel_wt / refvel_cfd;
    double R = r_rotor / reflength_wt;
    if (ti == tistart) {
        st[4] = 0.0;
        if (mode == 1) {
            double TSR = tsr_in;
            st[0] = TSR * uref / R;
            st[1] = tsr_in;
        } else if (mode == 2) {
            double TSR = TSR_max;
            st[0] = TSR * uref / R;
        } else {
            st[0] = angvel_fixed;
            st[1] = st[0] * R / uref;
        }
    } else {
        if (mode == 1) {
            st[1] = tsr_in;
            st[0] = st[1] * uref / R;
        } else if (mode == 0) {
            st[0] = angvel_fixed;
            st[1] = st[0] * R / uref;
        } else {
            double K = 0.5 * M_PI * pow(r_rotor, 2 + X_control) * CP_max / pow(TSR_max, X_control);
            double angvel_wt = st[0] / ratio_T;
            double UU = uref * ratio_V;
            if (!fixgen) st[3] = K * pow(angvel_wt, X_control - 1) * pow(UU, 3 - X_control);
            else st[3] = Tgen_fixed;
            double Torque_c = st[3];
            double Torque_a = st[2] * ratio_rho * pow(ratio_V, 2) * pow(ratio_L, 3);
            double Torque = Torque_a - Torque_c;
            double dt_wt = dt * ratio_T;
            double angvel_0 = st[0] / ratio_T;
            double angvel_1 = angvel_0 + dt_wt * Torque / J;
            st[0] = angvel_1 * ratio_T;
            st[1] = st[0] * R / uref;
        }
    }
    st[4] += st[0] * dt;
}
