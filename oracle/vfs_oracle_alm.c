/*
 * vfs_oracle_alm.c -- actuator-line delta kernels (rotor_model.c:5095-5393).
 * TEST INFRASTRUCTURE ONLY (see vfs_oracle.h).
 *
 * The truncated pi constants are the reference's own (3.14159265 in the
 * s3h/s4h kernels, 3.1415926 in dfunc_exp): matching them keeps parity at
 * round-off level (the difference to M_PI is ~1e-9 relative).
 */
#include "orc_internal.h"

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
