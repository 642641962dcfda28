/*
 * oracle_common.c -- source wavelet, step count and record predicate.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <math.h>

#include "oracle.h"

size_t o_ricker_length(double f, double dt)
{
    /* src.c:50-51: tcut = 1/f, length = ceil(tcut*2/dt) */
    double tcut = 1.0 / f;
    return (size_t)ceil(tcut * 2.0 / dt);
}

void o_compute_src_fct(size_t n, const double *A, const double *f, double dt,
                       const size_t *length, double *fct)
{
    /* src.c:45-64.  pi from atan (src.c:45), pi^2 (46). */
    const double pi = 4.0 * atan(1.0);
    const double pi2 = pi * pi;
    double *fct0 = fct;
    size_t len0 = length[0];
    size_t off = 0;
    for (size_t s = 0; s < n; ++s) {
        double *w = fct + off;
        const double tcut = 1.0 / f[s];
        double amax = 0.0;
        for (size_t i = 0; i < length[s]; ++i) {
            /* src.c:56-58: K = pi2 f f time time; A 2 pi2 f f (1-2K) e^-K */
            double time = i * dt - tcut;
            double K = pi2 * f[s] * f[s] * time * time;
            w[i] = A[s] * 2. * pi2 * f[s] * f[s] * (1.0 - 2.0 * K) * exp(-K);
            /* src.c:59: the maximum is taken over source 0's samples */
            if (i < len0) {
                double m = fabs(fct0[i]);
                amax = amax > m ? amax : m;
            }
        }
        /* src.c:61-63: source 0 is rescaled by A_s/Amax */
        for (size_t i = 0; i < len0; ++i)
            fct0[i] *= A[s] / amax;
        off += length[s];
    }
}

size_t o_nsteps(double duration, double dt)
{
    size_t nsteps = duration / dt;          /* e_iso.c:575 */
    if (nsteps * dt < duration) nsteps++;   /* e_iso.c:576 */
    return nsteps;
}

int o_record_due(size_t it, double dt, double t0, double rdt)
{
    /* e_iso.c:591, 643-644, 657-658 */
    double t = dt * (it + 1);
    double dt2 = t - t0 + 1.e-15;
    double rem = fmod(1.e6 * dt2, 1.e6 * rdt);
    return (fabs(dt2) < 0.99 * dt) ||
           (t > t0 && rdt > 0.0 && fabs(rem) < 1.e-8);
}
