/*
 * oracle_a3d.c -- FP64 restatement of a_iso_3d's setup and time loop.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Storage follows the reference: every array is full-grid, index
 * (i*ny2 + j)*nz2 + k with z fastest (a_iso_3d.c:608), and the six CPML
 * memory arrays are full-grid too (a_iso_3d.c:330-335).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

void o_cpml_tables(size_t n2, size_t np, double d, double dt, double vmax,
                   double pml_order, double Rc, double kappa_max,
                   double alpha_max, double *a, double *b, double *k,
                   double *a_h, double *b_h, double *k_h)
{
    /* interior values, a_iso_3d.c:500-523 */
    for (size_t n = 0; n < n2; ++n) { a[n] = 0.0; b[n] = 1.0; k[n] = 1.0; }
    for (size_t n = 0; n + 1 < n2; ++n) { a_h[n] = 0.0; b_h[n] = 1.0; k_h[n] = 1.0; }
    if (np == 0) return;
    /* a_iso_3d.c:528-534 */
    const double da = alpha_max / np;
    const double L = np * d;
    const double d0 = -(pml_order + 1.) * vmax * log(Rc) / (2. * L);
    /* a_iso_3d.c:536-576, one axis; node n mirrors to n2-1-n, half node n
     * mirrors to n2-2-n.  Writes are issued in the same order as the
     * reference so overlapping mirrors on tiny grids resolve the same way. */
    for (size_t n = 0; n < np; ++n) {
        double kn = 1. + (kappa_max - 1.) * pow(((np - n) * d / L), pml_order);
        double khn = 1. + (kappa_max - 1.) * pow(((np - n - 0.5) * d / L), pml_order);
        k[n] = k[n2 - 1 - n] = kn;
        k_h[n] = k_h[n2 - 2 - n] = khn;
        double alpha = n * da;
        double alpha_h = (n + 0.5) * da;
        double dn = d0 * pow(((np - n) * d / L), pml_order);
        double dn_h = d0 * pow(((np - n - 0.5) * d / L), pml_order);
        double bn = exp(-(dn / k[n] + alpha) * dt);
        b[n] = b[n2 - 1 - n] = bn;
        double bhn = exp(-(dn_h / k_h[n] + alpha_h) * dt);
        b_h[n] = b_h[n2 - 2 - n] = bhn;
        double an = dn * (b[n] - 1.) / (k[n] * (dn + k[n] * alpha));
        a[n] = a[n2 - 1 - n] = an;
        double ahn = dn_h * (b_h[n] - 1.) / (k_h[n] * (dn_h + k_h[n] * alpha_h));
        a_h[n] = a_h[n2 - 2 - n] = ahn;
    }
}

typedef struct {
    size_t nx2, ny2, nz2;
    double dx, dy, dz;
    const double *irho[3];
    const double *kdt;
    double *P, *v[3];
    double *mP[3], *mv[3];                 /* mem_dPd{x,y,z}, mem_dvd{x,y,z} */
    const double *ta[3], *tb[3], *tk[3];   /* node tables */
    const double *tah[3], *tbh[3], *tkh[3];/* half-node tables */
} a3d_state;

static inline size_t at(const a3d_state *s, size_t i, size_t j, size_t k)
{
    return (i * s->ny2 + j) * s->nz2 + k;
}

/* one velocity component sweep over x planes [i_begin, i_end),
 * a_iso_3d.c:605-651 (axis 0: x, 1: y, 2: z) */
static void sweep_velocity_rows(a3d_state *s, int axis, size_t i_begin, size_t i_end)
{
    const size_t ni = s->nx2 - (axis == 0), nj = s->ny2 - (axis == 1),
                 nk = s->nz2 - (axis == 2);
    if (i_end > ni) i_end = ni;
    const double h = axis == 0 ? s->dx : (axis == 1 ? s->dy : s->dz);
    const size_t step = axis == 0 ? s->ny2 * s->nz2 : (axis == 1 ? s->nz2 : 1);
    double *v = s->v[axis], *mem = s->mP[axis];
    const double *ir = s->irho[axis];
    const double *a = s->tah[axis], *b = s->tbh[axis], *kk = s->tkh[axis];
    for (size_t i = i_begin; i < i_end; ++i)
        for (size_t j = 0; j < nj; ++j)
            for (size_t k = 0; k < nk; ++k) {
                size_t c = at(s, i, j, k);
                size_t n = axis == 0 ? i : (axis == 1 ? j : k);
                double dP = (s->P[c + step] - s->P[c]) / h;
                mem[c] = b[n] * mem[c] + a[n] * dP;
                dP = dP / kk[n] + mem[c];
                v[c] -= ir[c] * dP;
            }
}

static void sweep_velocity(a3d_state *s, int axis) { sweep_velocity_rows(s, axis, 0, s->nx2); }

/* pressure sweep, a_iso_3d.c:659-683 */
static void sweep_pressure_rows(a3d_state *s, size_t i_begin, size_t i_end)
{
    const size_t sx = s->ny2 * s->nz2, sy = s->nz2;
    for (size_t i = i_begin; i < i_end; ++i)
        for (size_t j = 1; j < s->ny2; ++j)
            for (size_t k = 1; k < s->nz2; ++k) {
                size_t c = at(s, i, j, k);
                double dvdx = (s->v[0][c] - s->v[0][c - sx]) / s->dx;
                double dvdy = (s->v[1][c] - s->v[1][c - sy]) / s->dy;
                double dvdz = (s->v[2][c] - s->v[2][c - 1]) / s->dz;
                s->mv[0][c] = s->tb[0][i] * s->mv[0][c] + s->ta[0][i] * dvdx;
                dvdx = dvdx / s->tk[0][i] + s->mv[0][c];
                s->mv[1][c] = s->tb[1][j] * s->mv[1][c] + s->ta[1][j] * dvdy;
                dvdy = dvdy / s->tk[1][j] + s->mv[1][c];
                s->mv[2][c] = s->tb[2][k] * s->mv[2][c] + s->ta[2][k] * dvdz;
                dvdz = dvdz / s->tk[2][k] + s->mv[2][c];
                s->P[c] -= s->kdt[c] * (dvdx + dvdy + dvdz);
            }
}

int o_a3d_run(const o_a3d_params *p, const double *vp_in, const double *rho,
              size_t nsrc, const size_t *src_idx, const size_t *src_len,
              const double *src_fct, size_t nrec, const size_t *rec_idx,
              const int *rec_comp, size_t nsteps, double *traces,
              double *fields)
{
    if (p->nx == 0 || p->ny == 0 || p->nz == 0) return O_ERR_ARG;
    a3d_state s;
    s.nx2 = p->nx + 2 * p->np;
    s.ny2 = p->ny + 2 * p->np;
    s.nz2 = p->nz + 2 * p->np;
    s.dx = p->dx; s.dy = p->dy; s.dz = p->dz;
    const double dt = p->dt;
    const size_t nnn = s.nx2 * s.ny2 * s.nz2;

    double *buf = calloc(15 * nnn, sizeof(double));
    size_t n2[3] = {s.nx2, s.ny2, s.nz2};
    double *tab = calloc(6 * (s.nx2 + s.ny2 + s.nz2), sizeof(double));
    if (!buf || !tab) { free(buf); free(tab); return O_ERR_NOMEM; }
    double *irho[3] = {buf, buf + nnn, buf + 2 * nnn};
    double *kdt = buf + 3 * nnn;
    s.P = buf + 4 * nnn;
    for (int d = 0; d < 3; ++d) {
        s.v[d] = buf + (5 + d) * nnn;
        s.mP[d] = buf + (8 + d) * nnn;
        s.mv[d] = buf + (11 + d) * nnn;
        s.irho[d] = irho[d];
    }
    s.kdt = kdt;

    /* a_iso_3d.c:368-374: vpmax over the first nx*nz padded entries */
    double vpmax = 0.0;
    for (size_t i = 0; i < p->nx; ++i)
        for (size_t j = 0; j < p->nz; ++j) {
            size_t ind = i * p->nz + j;
            vpmax = vp_in[ind] > vpmax ? vp_in[ind] : vpmax;
        }
    /* a_iso_3d.c:379-405: irho_d = dt*2/(rho_a + rho_b) */
    for (size_t i = 0; i + 1 < s.nx2; ++i)
        for (size_t j = 0; j < s.ny2; ++j)
            for (size_t k = 0; k < s.nz2; ++k) {
                size_t c = at(&s, i, j, k);
                irho[0][c] = dt * 2. / (rho[c] + rho[at(&s, i + 1, j, k)]);
            }
    for (size_t i = 0; i < s.nx2; ++i)
        for (size_t j = 0; j + 1 < s.ny2; ++j)
            for (size_t k = 0; k < s.nz2; ++k) {
                size_t c = at(&s, i, j, k);
                irho[1][c] = dt * 2. / (rho[c] + rho[at(&s, i, j + 1, k)]);
            }
    for (size_t i = 0; i < s.nx2; ++i)
        for (size_t j = 0; j < s.ny2; ++j)
            for (size_t k = 0; k + 1 < s.nz2; ++k) {
                size_t c = at(&s, i, j, k);
                irho[2][c] = dt * 2. / (rho[c] + rho[c + 1]);
            }
    /* a_iso_3d.c:407-409: K*dt = rho vp^2 dt */
    for (size_t n = 0; n < nnn; ++n) kdt[n] = rho[n] * vp_in[n] * vp_in[n] * dt;
    /* a_iso_3d.c:411-417 */
    if (dt * vpmax * sqrt(1. / (p->dx * p->dx) + 1. / (p->dy * p->dy) +
                          1. / (p->dz * p->dz)) > 1.) {
        free(buf); free(tab);
        return O_ERR_CFL;
    }
    /* a_iso_3d.c:525-527 */
    const double pi = 4.0 * atan(1.0);
    double alpha_max = pi * p->f0;
    if (p->alpha_max == 0) alpha_max = 0.0;
    double *tp = tab;
    const double dd[3] = {p->dx, p->dy, p->dz};
    for (int d = 0; d < 3; ++d) {
        double *a = tp, *b = tp + n2[d], *k = tp + 2 * n2[d];
        double *ah = tp + 3 * n2[d], *bh = tp + 4 * n2[d], *kh = tp + 5 * n2[d];
        tp += 6 * n2[d];
        o_cpml_tables(n2[d], p->np, dd[d], dt, p->vmax, p->pml_order, p->Rc,
                      p->kappa_max, alpha_max, a, b, k, ah, bh, kh);
        s.ta[d] = a; s.tb[d] = b; s.tk[d] = k;
        s.tah[d] = ah; s.tbh[d] = bh; s.tkh[d] = kh;
    }

    const int nth = p->nthreads > 0 ? p->nthreads : 1;
    const int pth = p->pressure_threads > 0 ? p->pressure_threads : 1;
    const int sth = p->split_threads;
    (void)nth; (void)pth; (void)sth;
    for (size_t it = 0; it < nsteps; ++it) {
        if (sth > 0) {
            /* every sweep split over x planes: each cell's arithmetic is the
             * same as in the sections below, so the values are identical */
            for (int axis = 0; axis < 3; ++axis) {
#pragma omp parallel for num_threads(sth) schedule(static)
                for (size_t i = 0; i < s.nx2; ++i) sweep_velocity_rows(&s, axis, i, i + 1);
            }
#pragma omp parallel for num_threads(sth) schedule(static)
            for (size_t i = 1; i < s.nx2; ++i) sweep_pressure_rows(&s, i, i + 1);
        } else {
            /* velocity phase: three independent sections (a_iso_3d.c:599-652) */
#pragma omp parallel sections num_threads(nth)
            {
#pragma omp section
                sweep_velocity(&s, 0);
#pragma omp section
                sweep_velocity(&s, 1);
#pragma omp section
                sweep_velocity(&s, 2);
            }
            /* pressure phase: one section in the reference (a_iso_3d.c:655-684);
             * pressure_threads > 1 splits it over x rows (same arithmetic). */
            if (pth <= 1) {
                sweep_pressure_rows(&s, 1, s.nx2);
            } else {
#pragma omp parallel for num_threads(pth) schedule(static)
                for (size_t i = 1; i < s.nx2; ++i) sweep_pressure_rows(&s, i, i + 1);
            }
        }
        /* source, a_iso_3d.c:689-693 */
        size_t off = 0;
        for (size_t n = 0; n < nsrc; ++n) {
            if (it < src_len[n]) s.P[src_idx[n]] += dt * src_fct[off + it] * 1.e6;
            off += src_len[n];
        }
        /* records: field value after the step, a_iso_3d.c:700-723 */
        if (traces) {
            for (size_t r = 0; r < nrec; ++r) {
                const double *f = rec_comp[r] == O_VX ? s.v[0]
                                : rec_comp[r] == O_VY ? s.v[1]
                                : rec_comp[r] == O_VZ ? s.v[2] : s.P;
                traces[it * nrec + r] = f[rec_idx[r]];
            }
        }
    }
    if (fields) {
        memcpy(fields, s.P, nnn * sizeof(double));
        for (int d = 0; d < 3; ++d)
            memcpy(fields + (1 + d) * nnn, s.v[d], nnn * sizeof(double));
    }
    free(buf);
    free(tab);
    return O_OK;
}
