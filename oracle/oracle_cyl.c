/*
 * oracle_cyl.c -- FP64 restatement of e_cyl_iso, the 2.5D elastic program
 * in cylindrical coordinates (vertical symmetry axis at r = 0) with CPML on
 * the outer r side and at both z ends.  TEST INFRASTRUCTURE: only tests/,
 * the smoke check and the benchmark's CPU leg call it, as the checker.
 *
 * Follows, in the reference's operation order (no FMA contraction):
 *   material and CFL          e_cyl_iso.c:342-434
 *   CPML tables               pml.c:2107-2180 (compute_cpml_cyl)
 *   time loop                 e_cyl_iso.c:601-733 (n > 0) and 736-866 (n = 0)
 *   velocity updates          propagate.c:2706-2997 (n = 0), 3312-3693 (n > 0)
 *   stress updates            propagate.c:2999-3310 (n = 0), 3695-4088 (n > 0;
 *                             tau_rz uses the n = 0 routine in both modes,
 *                             e_cyl_iso.c:677)
 *   force sources             src.c:186-200 (add_force_src_cyl)
 *   Kurkjian source           src.c:202-238
 *   div / curl records        propagate.c:4090-4119
 *
 * Layout: every field and material array is nx2 x nz2 (index i*nz2 + j,
 * r = i*dr the row, z the fast axis), nx2 = nx + np (no layer at the axis),
 * nz2 = nz + 2 np.  The reference keeps its memory variables in compact
 * arrays filled by running counters; each slot belongs to one cell of one
 * update, so here every memory variable is a full nx2 x nz2 array indexed
 * by the cell (the values are the same).
 *
 * Staggering (e_cyl_iso.c:84-101): tau_rz at (i, j); v_z, tau_thetaz at
 * (i+1/2, j); v_r, tau_rtheta at (i, j+1/2); v_theta, tau_rr, tau_thetatheta,
 * tau_zz at (i+1/2, j+1/2).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* ------------------------------------------------------------ set-up */
int o_cyl_material(size_t nx2, size_t nz2, const double *vp, const double *vs, const double *rho,
                   double *lij, double *l2mij, double *mu, double *mui, double *muj, double *bi,
                   double *bj, double *bij, double *vpmax_out)
{
    /* e_cyl_iso.c:342-415; returns 0, or 1 + the first cell whose value is
     * not finite (the reference aborts there, 417-452) */
    const size_t N = nx2 * nz2;
    double *m = malloc(N * sizeof(double));
    if (!m) return -1;
    double vpmax = 0.0;
    for (size_t i = 0; i < nx2; ++i)
        for (size_t j = 0; j < nz2; ++j) {
            const size_t c = i * nz2 + j;
            vpmax = vp[c] > vpmax ? vp[c] : vpmax;
            lij[c] = rho[c] * (vp[c] * vp[c] - 2. * vs[c] * vs[c]);
            m[c] = rho[c] * vs[c] * vs[c];
            l2mij[c] = lij[c] + 2.0 * m[c];
            bij[c] = 1. / rho[c];
        }
    /* harmonic mu and arithmetic rho between z neighbours (j-1, j) */
    for (size_t i = 0; i < nx2; ++i) {
        const size_t r = i * nz2;
        muj[r] = m[r];
        bj[r] = bij[r];
        for (size_t j = 1; j < nz2; ++j) {
            const double a = m[r + j], b = m[r + j - 1];
            muj[r + j] = (a == 0.0 && b == 0.0) ? 0.0 : 2. * a * b / (a + b);
            bj[r + j] = 1. / ((rho[r + j] + rho[r + j - 1]) / 2.0);
        }
    }
    /* the same between r neighbours (i-1, i) */
    for (size_t j = 0; j < nz2; ++j) {
        mui[j] = m[j];
        bi[j] = 1. / rho[j];
        for (size_t i = 1; i < nx2; ++i) {
            const double a = m[i * nz2 + j], b = m[(i - 1) * nz2 + j];
            mui[i * nz2 + j] = (a == 0.0 && b == 0.0) ? 0.0 : 2. * a * b / (a + b);
            bi[i * nz2 + j] = 1. / ((rho[i * nz2 + j] + rho[(i - 1) * nz2 + j]) / 2.0);
        }
    }
    /* mu at (i, j): two-point harmonic means on the first row / column,
     * four-point elsewhere */
    mu[0] = m[0];
    for (size_t j = 1; j < nz2; ++j) {
        const double a = m[j], b = m[j - 1];
        mu[j] = (a == 0.0 && b == 0.0) ? 0.0 : 2. * a * b / (a + b);
    }
    for (size_t i = 1; i < nx2; ++i) {
        const double a0 = m[i * nz2], b0 = m[(i - 1) * nz2];
        mu[i * nz2] = (a0 == 0.0 && b0 == 0.0) ? 0.0 : 2. * a0 * b0 / (a0 + b0);
        for (size_t j = 1; j < nz2; ++j) {
            const double p = m[i * nz2 + j], q = m[i * nz2 + j - 1], s = m[(i - 1) * nz2 + j],
                         t = m[(i - 1) * nz2 + j - 1];
            if (q * s * t == 0.0 && p * s * t == 0.0 && p * q * t == 0.0 && p * q * s == 0.0)
                mu[i * nz2 + j] = 0.0;
            else
                mu[i * nz2 + j] = 4. * p * q * s * t / (q * s * t + p * s * t + p * q * t + p * q * s);
        }
    }
    free(m);
    *vpmax_out = vpmax;
    for (size_t c = 0; c < N; ++c)
        if (!isfinite(mu[c]) || !isfinite(mui[c]) || !isfinite(muj[c]) || !isfinite(bi[c]) ||
            !isfinite(bj[c]) || !isfinite(bij[c]) || !isfinite(lij[c]) || !isfinite(l2mij[c]))
            return (int)(1 + (c < 0x7ffffffe ? c : 0x7ffffffe));
    return 0;
}

void o_cyl_cpml(size_t np, double dx, double dz, double dt, double vmax, double order, double Rc,
                double kappa_max, double alpha_max, double *t)
{
    /* pml.c:2107-2164.  t holds 9 tables of np entries (ik_r, ik_z, ikh_z,
     * b_r, c_r, b_z, c_z, bh_z, ch_z), then 6 of np+1 (ikH_r, ikH_z, bH_r,
     * cH_r, bH_z, cH_z). */
    double *ik_r = t, *ik_z = t + np, *ikh_z = t + 2 * np, *b_r = t + 3 * np, *c_r = t + 4 * np,
           *b_z = t + 5 * np, *c_z = t + 6 * np, *bh_z = t + 7 * np, *ch_z = t + 8 * np;
    double *H = t + 9 * np;
    double *ikH_r = H, *ikH_z = H + (np + 1), *bH_r = H + 2 * (np + 1), *cH_r = H + 3 * (np + 1),
           *bH_z = H + 4 * (np + 1), *cH_z = H + 5 * (np + 1);
    const double Lx = np * dx, Lz = np * dz;
    const double d0x = -(order + 1.) * vmax * log(Rc) / (2. * Lx);
    const double d0z = -(order + 1.) * vmax * log(Rc) / (2. * Lz);
    const double da = alpha_max / np, da2 = alpha_max / (np + 1);
    for (size_t n = 0; n < np; ++n) {
        const double k_r = 1. + (kappa_max - 1.) * pow(((np - n) * dx / Lx), order);
        const double k_z = 1. + (kappa_max - 1.) * pow(((np - n) * dz / Lz), order);
        ik_r[n] = 1. / k_r;
        ik_z[n] = 1. / k_z;
        /* the half-node kappa along z is formed with dx / Lx (pml.c:2128) */
        const double kh_z = 1. + (kappa_max - 1.) * pow(((np - n - 0.5) * dx / Lx), order);
        ikh_z[n] = 1. / kh_z;
        const double d_x = d0x * pow(((np - n) * dx / Lx), order);
        const double alpha_x = n * da;
        b_r[n] = exp(-(d_x / k_r + alpha_x) * dt);
        c_r[n] = d_x * (b_r[n] - 1.) / (k_r * (d_x + k_r * alpha_x));
        const double d_z = d0z * pow(((np - n) * dz / Lz), order);
        const double alpha_z = n * da;
        const double dh_z = d0z * pow(((np - n - 0.5) * dz / Lz), order);
        const double alphah_z = (n + 0.5) * da;
        b_z[n] = exp(-(d_z / k_z + alpha_z) * dt);
        c_z[n] = d_z * (b_z[n] - 1.) / (k_z * (d_z + k_z * alpha_z));
        bh_z[n] = exp(-(dh_z / kh_z + alphah_z) * dt);
        ch_z[n] = dh_z * (bh_z[n] - 1.) / (kh_z * (dh_z + kh_z * alphah_z));
    }
    for (size_t n = 0; n < np + 1; ++n) {
        const double kH_r = 1. + (kappa_max - 1.) * pow(((n + 0.5) * dx / Lx), order);
        const double kH_z = 1. + (kappa_max - 1.) * pow(((n + 0.5) * dz / Lz), order);
        ikH_r[n] = 1. / kH_r;
        ikH_z[n] = 1. / kH_z;
        const double dH_x = d0x * pow(((n + 0.5) * dx / Lx), order);
        const double alphaH_x = (np - n - 0.5) * da2;
        bH_r[n] = exp(-(dH_x / kH_r + alphaH_x) * dt);
        cH_r[n] = dH_x * (bH_r[n] - 1.) / (kH_r * (dH_x + kH_r * alphaH_x));
        const double dH_z = d0z * pow(((n + 0.5) * dz / Lz), order);
        const double alphaH_z = (np - n - 0.5) * da2;
        bH_z[n] = exp(-(dH_z / kH_z + alphaH_z) * dt);
        cH_z[n] = dH_z * (bH_z[n] - 1.) / (kH_z * (dH_z + kH_z * alphaH_z));
    }
}

void o_kurkjian_fct(double lij_src, double dx, double dz, double x, int n, double A, double f,
                    double dt, size_t length, double *fct)
{
    /* src.c:202-229: unit Ricker normalised to peak 1, times A and the
     * Kurkjian factor lambda/(2 pi) eps_m 2/(dx^2 dz) (x/dx)^n */
    const double pi = 4.0 * atan(1.0);
    const double pi2 = pi * pi;
    const double eps = n > 0 ? 2.0 : 1.0;
    const double term = lij_src / (2. * pi) * eps * 2. / (dx * dx * dz) * pow(x / dx, n);
    const double tcut = 1.0 / f;
    double amax = 0.0;
    for (size_t i = 0; i < length; ++i) {
        const double time = i * dt - tcut;
        const double K = pi2 * f * f * time * time;
        fct[i] = 2. * pi2 * f * f * (1.0 - 2.0 * K) * exp(-K);
        amax = amax > fabs(fct[i]) ? amax : fabs(fct[i]);
    }
    for (size_t i = 0; i < length; ++i) fct[i] *= A / amax * term;
}

/* ------------------------------------------------------------ time loop */
enum { C_VR, C_VT, C_VZ, C_TRR, C_TTT, C_TZZ, C_TRZ, C_TRT, C_TTZ, C_NF };
/* memory variables, one full array each */
enum {
    M_DTRR_DR, M_TRR_R1, M_TRT_R1, M_TTT_R1, M_DTRZ_DZ, M_TRT_R2, M_DTRT_DR, M_TTT_R2, M_DTTZ_DZ,
    M_DTZZ_DZ, M_DTRZ_DR, M_TRZ_R, M_TTZ_R, M_DVR_DR, M_VR_R1, M_VT_R1, M_DVZ_DZ, M_DVR_DZ,
    M_DVZ_DR, M_DVT_DR, M_VR_R2, M_VT_R2, M_DVT_DZ, M_VZ_R, M_NM
};

typedef struct {
    size_t nx, nz, np, nx2, nz2;
    double dx, dz;
    int n;
    const double *ik_r, *ik_z, *ikh_z, *b_r, *c_r, *b_z, *c_z, *bh_z, *ch_z;
    const double *ikH_r, *ikH_z, *bH_r, *cH_r, *bH_z, *cH_z;
} Cyl;

/* CPML: memory variable update and the stretched derivative */
static inline double cpml(double *m, double b, double c, double ik, double d)
{
    *m = b * *m + c * d;
    return ik * d + *m;
}

/* coefficients (b, c, 1/kappa) of a row or column */
typedef struct {
    int on;
    double b, c, ik;
} Coef;

/* z coefficients at half nodes (fields at j+1/2: v_r, v_theta, taa; and
 * tau_rz, which uses them at its node rows): bh_z[j] on top, bH_z[j -
 * (nz+np-1)] at the bottom (top/bottom boundaries of the update loops as
 * the reference draws them) */
static Coef zc_half(const Cyl *g, size_t j, size_t j_bot)
{
    Coef k = {0, 0, 0, 1};
    if (j < g->np) { k.on = 1; k.b = g->bh_z[j]; k.c = g->ch_z[j]; k.ik = g->ikh_z[j]; }
    else if (j >= j_bot) {
        const size_t jj = j - j_bot;
        k.on = 1; k.b = g->bH_z[jj]; k.c = g->cH_z[jj]; k.ik = g->ikH_z[jj];
    }
    return k;
}
/* z coefficients at nodes (v_z, tau_thetaz): b_z[j] on top, b_z[np-1-jj]
 * at the bottom from j = nz+np */
static Coef zc_node(const Cyl *g, size_t j)
{
    Coef k = {0, 0, 0, 1};
    if (j < g->np) { k.on = 1; k.b = g->b_z[j]; k.c = g->c_z[j]; k.ik = g->ik_z[j]; }
    else if (j >= g->nz + g->np) {
        const size_t q = g->np - 1 - (j - (g->nz + g->np));
        k.on = 1; k.b = g->b_z[q]; k.c = g->c_z[q]; k.ik = g->ik_z[q];
    }
    return k;
}
/* r coefficients of rows at r = i dr (v_r, tau_rtheta): b_r[np-1-ii] from
 * i = nx */
static Coef rc_node(const Cyl *g, size_t i)
{
    Coef k = {0, 0, 0, 1};
    if (i >= g->nx) {
        const size_t q = g->np - 1 - (i - g->nx);
        k.on = 1; k.b = g->b_r[q]; k.c = g->c_r[q]; k.ik = g->ik_r[q];
    }
    return k;
}
/* r coefficients bH_r[i - i0] of rows from i0 (i0 = nx-1 for fields at
 * i+1/2; tau_rz uses them from i0 = nx) */
static Coef rc_half(const Cyl *g, size_t i, size_t i0)
{
    Coef k = {0, 0, 0, 1};
    if (i >= i0) {
        const size_t ii = i - i0;
        k.on = 1; k.b = g->bH_r[ii]; k.c = g->cH_r[ii]; k.ik = g->ikH_r[ii];
    }
    return k;
}

#define F(f, i, j) (V[f][(i) * nz2 + (j)])
#define MEM(q, c) (&M[q][c])

static void velocity(const Cyl *g, double *const *V, double *const *M, const double *bj,
                     const double *bi, const double *bij, double dt)
{
    const size_t nx = g->nx, np = g->np, nx2 = g->nx2, nz2 = g->nz2, nz = g->nz;
    const double dx = g->dx, dz = g->dz;
    const int n = g->n;
    const double dn = (double)n;
    /* ---- v_r, rows i = 1 .. nx2-1 (and i = 0 for odd n) */
    for (size_t i = (n % 2) ? 0 : 1; i < nx2; ++i) {
        const Coef kr = rc_node(g, i);
        for (size_t j = 0; j + 1 < nz2; ++j) {
            const size_t c = i * nz2 + j;
            const Coef kz = zc_half(g, j, nz + np - 1);
            double ddz = (F(C_TRZ, i, j + 1) - F(C_TRZ, i, j)) / dz;
            if (kz.on) ddz = cpml(MEM(M_DTRZ_DZ, c), kz.b, kz.c, kz.ik, ddz);
            if (i == 0) {
                /* odd mode on the axis, propagate.c:3321-3355 */
                F(C_VR, i, j) += dt * bj[c] * ((F(C_TRR, i, j) - -F(C_TRR, i, j)) / dx + ddz);
                continue;
            }
            const double twoidx = 2. * i * dx, idx_ = i * dx;
            double ddr1 = (F(C_TRR, i, j) - F(C_TRR, i - 1, j)) / dx;
            double dr1 = (F(C_TRR, i, j) + F(C_TRR, i - 1, j)) / twoidx;
            double dr3 = (F(C_TTT, i, j) + F(C_TTT, i - 1, j)) / twoidx;
            if (n == 0) {
                if (kr.on) {
                    ddr1 = cpml(MEM(M_DTRR_DR, c), kr.b, kr.c, kr.ik, ddr1);
                    dr1 = cpml(MEM(M_TRR_R1, c), kr.b, kr.c, kr.ik, dr1);
                    dr3 = cpml(MEM(M_TTT_R1, c), kr.b, kr.c, kr.ik, dr3);
                }
                F(C_VR, i, j) += dt * bj[c] * (ddr1 + dr1 - dr3 + ddz);
            } else if (kr.on) {
                ddr1 = cpml(MEM(M_DTRR_DR, c), kr.b, kr.c, kr.ik, ddr1);
                dr1 = cpml(MEM(M_TRR_R1, c), kr.b, kr.c, kr.ik, dr1);
                double dr2 = F(C_TRT, i, j) / idx_;
                dr2 = cpml(MEM(M_TRT_R1, c), kr.b, kr.c, kr.ik, dr2);
                dr3 = cpml(MEM(M_TTT_R1, c), kr.b, kr.c, kr.ik, dr3);
                F(C_VR, i, j) += dt * bj[c] * (ddr1 + dr1 + dn * dr2 - dr3 + ddz);
            } else {
                F(C_VR, i, j) += dt * bj[c] * (ddr1 + dr1 + dn * F(C_TRT, i, j) / idx_ - dr3 + ddz);
            }
        }
    }
    /* ---- v_theta, rows i = 0 .. nx2-2 */
    for (size_t i = 0; i + 1 < nx2; ++i) {
        const Coef kr = rc_half(g, i, nx - 1);
        const double ihdx = (i + 0.5) * dx;
        for (size_t j = 0; j + 1 < nz2; ++j) {
            const size_t c = i * nz2 + j;
            const Coef kz = zc_half(g, j, nz + np - 1);
            double ddz = (F(C_TTZ, i, j + 1) - F(C_TTZ, i, j)) / dz;
            if (kz.on) ddz = cpml(MEM(M_DTTZ_DZ, c), kz.b, kz.c, kz.ik, ddz);
            double dr1 = (F(C_TRT, i + 1, j) + F(C_TRT, i, j)) / ihdx;
            double ddr = (F(C_TRT, i + 1, j) - F(C_TRT, i, j)) / dx;
            if (kr.on) {
                dr1 = cpml(MEM(M_TRT_R2, c), kr.b, kr.c, kr.ik, dr1);
                ddr = cpml(MEM(M_DTRT_DR, c), kr.b, kr.c, kr.ik, ddr);
            }
            if (n == 0) {
                F(C_VT, i, j) += dt * bij[c] * (dr1 + ddr + ddz);
            } else if (kr.on) {
                double dr2 = F(C_TTT, i, j) / ihdx;
                dr2 = cpml(MEM(M_TTT_R2, c), kr.b, kr.c, kr.ik, dr2);
                F(C_VT, i, j) += dt * bij[c] * (dr1 + ddr - dn * dr2 + ddz);
            } else {
                F(C_VT, i, j) += dt * bij[c] * (dr1 + ddr - dn * F(C_TTT, i, j) / ihdx + ddz);
            }
        }
    }
    /* ---- v_z, rows i = 0 .. nx2-2, j = 1 .. nz2-1 */
    for (size_t i = 0; i + 1 < nx2; ++i) {
        const Coef kr = rc_half(g, i, nx - 1);
        const double ihdx = (i + 0.5) * dx, two_ihdx = 2. * (i + 0.5) * dx;
        for (size_t j = 1; j < nz2; ++j) {
            const size_t c = i * nz2 + j;
            const Coef kz = zc_node(g, j);
            double ddz = (F(C_TZZ, i, j) - F(C_TZZ, i, j - 1)) / dz;
            if (kz.on) ddz = cpml(MEM(M_DTZZ_DZ, c), kz.b, kz.c, kz.ik, ddz);
            double ddr = (F(C_TRZ, i + 1, j) - F(C_TRZ, i, j)) / dx;
            double dr1 = (F(C_TRZ, i + 1, j) + F(C_TRZ, i, j)) / two_ihdx;
            if (kr.on) {
                ddr = cpml(MEM(M_DTRZ_DR, c), kr.b, kr.c, kr.ik, ddr);
                dr1 = cpml(MEM(M_TRZ_R, c), kr.b, kr.c, kr.ik, dr1);
            }
            if (n == 0) {
                F(C_VZ, i, j) += dt * bi[c] * (ddz + ddr + dr1);
            } else if (kr.on) {
                double dr2 = F(C_TTZ, i, j) / ihdx;
                dr2 = cpml(MEM(M_TTZ_R, c), kr.b, kr.c, kr.ik, dr2);
                F(C_VZ, i, j) += dt * bi[c] * (ddz + ddr + dr1 + dn * dr2);
            } else {
                F(C_VZ, i, j) += dt * bi[c] * (ddz + ddr + dr1 + dn * F(C_TTZ, i, j) / ihdx);
            }
        }
    }
}

static void stress(const Cyl *g, double *const *V, double *const *M, const double *lij,
                   const double *l2mij, const double *mu, const double *mui, const double *muj,
                   double dt)
{
    const size_t nx = g->nx, np = g->np, nx2 = g->nx2, nz2 = g->nz2, nz = g->nz;
    const double dx = g->dx, dz = g->dz;
    const int n = g->n;
    const double dn = (double)n;
    /* ---- tau_rr, tau_thetatheta, tau_zz: rows 0 .. nx2-2, j = 0 .. nz2-2 */
    for (size_t i = 0; i + 1 < nx2; ++i) {
        const Coef kr = rc_half(g, i, nx - 1);
        const double ihdx = (i + 0.5) * dx, two_ihdx = 2. * (i + 0.5) * dx;
        for (size_t j = 0; j + 1 < nz2; ++j) {
            const size_t c = i * nz2 + j;
            const Coef kz = zc_half(g, j, nz + np - 1);
            double dvrdr = (F(C_VR, i + 1, j) - F(C_VR, i, j)) / dx;
            double vrr = (F(C_VR, i + 1, j) + F(C_VR, i, j)) / two_ihdx;
            double vtr = 0.0;
            if (n != 0) vtr = F(C_VT, i, j) / ihdx;
            double dvzdz = (F(C_VZ, i, j + 1) - F(C_VZ, i, j)) / dz;
            if (kr.on) {
                dvrdr = cpml(MEM(M_DVR_DR, c), kr.b, kr.c, kr.ik, dvrdr);
                vrr = cpml(MEM(M_VR_R1, c), kr.b, kr.c, kr.ik, vrr);
                if (n != 0) vtr = cpml(MEM(M_VT_R1, c), kr.b, kr.c, kr.ik, vtr);
            }
            if (kz.on) dvzdz = cpml(MEM(M_DVZ_DZ, c), kz.b, kz.c, kz.ik, dvzdz);
            if (n == 0) {
                F(C_TRR, i, j) += dt * l2mij[c] * dvrdr + dt * lij[c] * (vrr + dvzdz);
                F(C_TTT, i, j) += dt * l2mij[c] * vrr + dt * lij[c] * (dvrdr + dvzdz);
                F(C_TZZ, i, j) += dt * l2mij[c] * dvzdz + dt * lij[c] * (dvrdr + vrr);
            } else {
                F(C_TRR, i, j) += dt * l2mij[c] * dvrdr + dt * lij[c] * (vrr + dn * vtr + dvzdz);
                F(C_TTT, i, j) += dt * l2mij[c] * (vrr + dn * vtr) + dt * lij[c] * (dvrdr + dvzdz);
                F(C_TZZ, i, j) += dt * l2mij[c] * dvzdz + dt * lij[c] * (dvrdr + vrr + dn * vtr);
            }
        }
    }
    /* ---- tau_rz (both modes, propagate.c:3136-3213): rows 1 .. nx2-1,
     * j = 1 .. nz2-1; the z layers use the half-node tables and the r layer
     * bH_r from i = nx */
    for (size_t i = 1; i < nx2; ++i) {
        const Coef kr = rc_half(g, i, nx);
        for (size_t j = 1; j < nz2; ++j) {
            const size_t c = i * nz2 + j;
            const Coef kz = zc_half(g, j, nz + np);
            double ddz = (F(C_VR, i, j) - F(C_VR, i, j - 1)) / dz;
            double ddr = (F(C_VZ, i, j) - F(C_VZ, i - 1, j)) / dx;
            if (kz.on) ddz = cpml(MEM(M_DVR_DZ, c), kz.b, kz.c, kz.ik, ddz);
            if (kr.on) ddr = cpml(MEM(M_DVZ_DR, c), kr.b, kr.c, kr.ik, ddr);
            F(C_TRZ, i, j) += dt * mu[c] * (ddz + ddr);
        }
    }
    /* ---- tau_rtheta: rows 1 .. nx2-1 (and 0 for odd n), every j */
    for (size_t i = (n % 2) ? 0 : 1; i < nx2; ++i) {
        const Coef kr = rc_node(g, i);
        const double twoidx = 2. * i * dx, idx_ = i * dx;
        for (size_t j = 0; j < nz2; ++j) {
            const size_t c = i * nz2 + j;
            if (i == 0) {
                /* odd mode on the axis, propagate.c:3963-3972 */
                F(C_TRT, i, j) += dt * muj[c] * ((F(C_VT, i, j) - -F(C_VT, i, j)) / dx);
                continue;
            }
            double ddr = (F(C_VT, i, j) - F(C_VT, i - 1, j)) / dx;
            double vtr = (F(C_VT, i, j) + F(C_VT, i - 1, j)) / twoidx;
            if (n == 0) {
                if (kr.on) {
                    ddr = cpml(MEM(M_DVT_DR, c), kr.b, kr.c, kr.ik, ddr);
                    vtr = cpml(MEM(M_VT_R2, c), kr.b, kr.c, kr.ik, vtr);
                }
                F(C_TRT, i, j) += dt * muj[c] * (ddr - vtr);
            } else if (kr.on) {
                ddr = cpml(MEM(M_DVT_DR, c), kr.b, kr.c, kr.ik, ddr);
                double vrr = F(C_VR, i, j) / idx_;
                vrr = cpml(MEM(M_VR_R2, c), kr.b, kr.c, kr.ik, vrr);
                vtr = cpml(MEM(M_VT_R2, c), kr.b, kr.c, kr.ik, vtr);
                F(C_TRT, i, j) += dt * muj[c] * (ddr - dn * vrr - vtr);
            } else {
                F(C_TRT, i, j) += dt * muj[c] * (ddr - dn * F(C_VR, i, j) / idx_ - vtr);
            }
        }
    }
    /* ---- tau_thetaz: rows 0 .. nx2-2, j = 1 .. nz2-1 */
    for (size_t i = 0; i + 1 < nx2; ++i) {
        const Coef kr = rc_half(g, i, nx - 1);
        const double ihdx = (i + 0.5) * dx;
        for (size_t j = 1; j < nz2; ++j) {
            const size_t c = i * nz2 + j;
            const Coef kz = zc_node(g, j);
            const double dvt = F(C_VT, i, j) - F(C_VT, i, j - 1);
            if (n == 0) {
                /* propagate.c:3249-3309: the interior form divides the
                 * product dt mu (v - v) by dz, the layers scale dv/dz */
                if (kz.on) {
                    const double ddz = cpml(MEM(M_DVT_DZ, c), kz.b, kz.c, kz.ik, dvt / dz);
                    F(C_TTZ, i, j) += dt * mui[c] * ddz;
                } else {
                    F(C_TTZ, i, j) += dt * mui[c] * dvt / dz;
                }
                continue;
            }
            double ddz = dvt / dz;
            if (kz.on) ddz = cpml(MEM(M_DVT_DZ, c), kz.b, kz.c, kz.ik, ddz);
            if (kr.on) {
                double vzr = F(C_VZ, i, j) / ihdx;
                vzr = cpml(MEM(M_VZ_R, c), kr.b, kr.c, kr.ik, vzr);
                F(C_TTZ, i, j) += dt * mui[c] * (ddz - dn * vzr);
            } else {
                F(C_TTZ, i, j) += dt * mui[c] * (ddz - dn * F(C_VZ, i, j) / ihdx);
            }
        }
    }
}

static void div_curl(const Cyl *g, double *const *V, double *div, double *curl)
{
    /* propagate.c:4090-4119 (the divergence's first term is not divided
     * by dr there; kept) */
    const size_t nx2 = g->nx2, nz2 = g->nz2;
    const double dx = g->dx, dz = g->dz;
    const int n = g->n;
    if (div)
        for (size_t i = 0; i + 1 < nx2; ++i)
            for (size_t j = 0; j + 1 < nz2; ++j)
                div[i * nz2 + j] = ((i + 1) * F(C_VR, i + 1, j) - i * F(C_VR, i, j)) / (i + 0.5) +
                                   n * F(C_VT, i, j) / ((i + 0.5) * dx) +
                                   (F(C_VZ, i, j + 1) - F(C_VZ, i, j)) / dz;
    if (curl)
        for (size_t i = (n % 2) ? 0 : 1; i < nx2; ++i)
            for (size_t j = 1; j < nz2; ++j)
                curl[i * nz2 + j] = (F(C_VR, i, j) - F(C_VR, i, j - 1)) / dz -
                                    (i == 0 ? (F(C_VZ, i, j) - -F(C_VZ, i, j))
                                            : (F(C_VZ, i, j) - F(C_VZ, i - 1, j))) / dx;
}

#undef F
#undef MEM

int o_cyl_run(const OCylParams *p, const double *vp, const double *vs, const double *rho,
              size_t nsrc, const int *type, const size_t *idx, const size_t *len, const double *fct,
              size_t nrec, const size_t *ridx, const int *rcomp, size_t nsteps, double *traces,
              double *fields)
{
    const size_t np = p->np, nx2 = p->nx + np, nz2 = p->nz + 2 * np, N = nx2 * nz2;
    if (p->nx < 2 || p->nz < 2) return O_ERR_ARG;
    double *buf = calloc((size_t)(8 + C_NF + M_NM + 2) * N + 15 * (np + 1), sizeof(double));
    if (!buf) return O_ERR_NOMEM;
    double *lij = buf, *l2mij = buf + N, *mu = buf + 2 * N, *mui = buf + 3 * N, *muj = buf + 4 * N,
           *bi = buf + 5 * N, *bj = buf + 6 * N, *bij = buf + 7 * N;
    double *V[C_NF], *M[M_NM];
    for (int f = 0; f < C_NF; ++f) V[f] = buf + (size_t)(8 + f) * N;
    for (int q = 0; q < M_NM; ++q) M[q] = buf + (size_t)(8 + C_NF + q) * N;
    double *div = buf + (size_t)(8 + C_NF + M_NM) * N, *curl = div + N;
    double *tab = curl + N;
    double vpmax = 0.0;
    if (o_cyl_material(nx2, nz2, vp, vs, rho, lij, l2mij, mu, mui, muj, bi, bj, bij, &vpmax) != 0) {
        free(buf);
        return O_ERR_ARG;
    }
    if (p->dt > 0.5 * (p->dx + p->dz) / (sqrt(2.0) * vpmax)) {   /* e_cyl_iso.c:454-458 */
        free(buf);
        return O_ERR_CFL;
    }
    if (np > 0)
        o_cyl_cpml(np, p->dx, p->dz, p->dt, p->vmax, p->pml_order, p->Rc, p->kappa_max,
                   p->alpha_max, tab);
    Cyl g = {p->nx, p->nz, np, nx2, nz2, p->dx, p->dz, p->n,
             tab, tab + np, tab + 2 * np, tab + 3 * np, tab + 4 * np, tab + 5 * np, tab + 6 * np,
             tab + 7 * np, tab + 8 * np,
             tab + 9 * np, tab + 9 * np + (np + 1), tab + 9 * np + 2 * (np + 1),
             tab + 9 * np + 3 * (np + 1), tab + 9 * np + 4 * (np + 1), tab + 9 * np + 5 * (np + 1)};
    size_t *off = malloc((nsrc + 1) * sizeof(size_t));
    off[0] = 0;
    for (size_t s = 0; s < nsrc; ++s) off[s + 1] = off[s] + len[s];
    const double dt = p->dt;
    for (size_t it = 0; it < nsteps; ++it) {
        velocity(&g, V, M, bj, bi, bij, dt);
        /* add_force_src_cyl, src.c:186-200 */
        for (size_t s = 0; s < nsrc; ++s) {
            if (it >= len[s]) continue;
            const double a = fct[off[s] + it];
            const size_t c = idx[s];
            if (type[s] == O_SRC_FR || type[s] == O_SRC_BULK) V[C_VR][c] += dt * bj[c] * a;
            if (type[s] == O_SRC_FT) V[C_VT][c] += dt * bij[c] * a;
            if (type[s] == O_SRC_FZ || type[s] == O_SRC_BULK) V[C_VZ][c] += dt * bi[c] * a;
        }
        stress(&g, V, M, lij, l2mij, mu, mui, muj, dt);
        /* add_kurkjian_src, src.c:231-238: a source set whose first entry
         * is Kurkjian injects that entry only */
        if (nsrc && type[0] == O_SRC_KURKJIAN && it < len[0]) {
            V[C_TRR][idx[0]] += dt * fct[it];
            V[C_TZZ][idx[0]] += dt * fct[it];
        }
        if (traces && nrec) {
            int need_div = 0, need_curl = 0;
            for (size_t r = 0; r < nrec; ++r) {
                need_div |= rcomp[r] == O_DIV;
                need_curl |= rcomp[r] == O_CURL;
            }
            if (need_div || need_curl) div_curl(&g, V, need_div ? div : NULL, need_curl ? curl : NULL);
            for (size_t r = 0; r < nrec; ++r) {
                const double *src;
                switch (rcomp[r]) {
                case O_VR: src = V[C_VR]; break;
                case O_VT: src = V[C_VT]; break;
                case O_VZ: src = V[C_VZ]; break;
                case O_TRR: src = V[C_TRR]; break;
                case O_TRT: src = V[C_TRT]; break;
                case O_TRZ: src = V[C_TRZ]; break;
                case O_TTT: src = V[C_TTT]; break;
                case O_TTZ: src = V[C_TTZ]; break;
                case O_TZZ: src = V[C_TZZ]; break;
                case O_DIV: src = div; break;
                case O_CURL: src = curl; break;
                default: src = NULL;
                }
                traces[it * nrec + r] = src ? src[ridx[r]] : 0.0;
            }
        }
    }
    if (fields)
        for (int f = 0; f < C_NF; ++f) memcpy(fields + (size_t)f * N, V[f], N * sizeof(double));
    free(off);
    free(buf);
    return O_OK;
}
