/*
 * oracle_e2d.c -- FP64 restatement of e_iso's setup and time loop.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Storage follows the reference: the five interior fields are nx*nz arrays
 * (index i*nz+j, z fastest; e_iso.c:384-388) and the ten split PML fields
 * live in one array each, partitioned into the eight regions of pml.c
 * (upper-left corner, left strip, lower-left corner, upper strip, lower
 * strip, upper-right corner, right strip, lower-right corner; offsets as
 * in pml.c:54, 98, 133, 177, 213, 248, 290, 325).
 *
 * The per-region behaviour of every split component (damped or not, which
 * damping table and index, which interior material value) is written out
 * as tables below, each row read from the pml.c loop it cites.  Neighbour
 * values are fetched by padded coordinate: an interior cell gives its
 * interior value, a PML cell gives x+z of its split pair, and a cell off
 * the padded grid gives 0 (pml.c:58, 79, 480, ...).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

void o_e2d_material(size_t nx, size_t nz, const double *vp, const double *vs,
                    const double *rho, double *la, double *mu, double *l2m,
                    double *b, double *bm)
{
    const size_t n = nx * nz;
    double *la_t = malloc(n * sizeof(double));
    double *mu_t = malloc(n * sizeof(double));
    double *l2m_t = malloc(n * sizeof(double));
    /* e_iso.c:424-432 */
    for (size_t c = 0; c < n; ++c) {
        la_t[c] = rho[c] * (vp[c] * vp[c] - 2. * vs[c] * vs[c]);
        mu_t[c] = rho[c] * vs[c] * vs[c];
        l2m_t[c] = la_t[c] + 2. * mu_t[c];
        b[c] = 1. / rho[c];
    }
    /* e_iso.c:438-443: mu averaged along z */
    for (size_t i = 0; i < nx; ++i) {
        for (size_t j = 0; j + 1 < nz; ++j)
            mu[i * nz + j] = (mu_t[i * nz + j] + mu_t[i * nz + j + 1]) / 2.0;
        mu[i * nz + nz - 1] = mu_t[i * nz + nz - 1];
    }
    /* e_iso.c:444-451: la, l2m averaged along x */
    for (size_t j = 0; j < nz; ++j) {
        for (size_t i = 0; i + 1 < nx; ++i) {
            la[i * nz + j] = (la_t[i * nz + j] + la_t[(i + 1) * nz + j]) / 2.0;
            l2m[i * nz + j] = (l2m_t[i * nz + j] + l2m_t[(i + 1) * nz + j]) / 2.0;
        }
        la[(nx - 1) * nz + j] = la_t[(nx - 1) * nz + j];
        l2m[(nx - 1) * nz + j] = l2m_t[(nx - 1) * nz + j];
    }
    /* e_iso.c:452-462: bm, four-point average of 1/rho */
    for (size_t i = 0; i + 1 < nx; ++i) {
        for (size_t j = 0; j + 1 < nz; ++j)
            bm[i * nz + j] = (b[i * nz + j] + b[(i + 1) * nz + j] +
                              b[i * nz + j + 1] + b[(i + 1) * nz + j + 1]) / 4.0;
        bm[i * nz + nz - 1] = (b[i * nz + nz - 1] + b[(i + 1) * nz + nz - 1]) / 2.0;
    }
    for (size_t j = 0; j + 1 < nz; ++j)
        bm[(nx - 1) * nz + j] = (b[(nx - 1) * nz + j] + b[(nx - 1) * nz + j + 1]) / 2.0;
    bm[n - 1] = b[n - 1];
    free(la_t); free(mu_t); free(l2m_t);
}

void o_e2d_damping(size_t npml, double h, double vp_mean, double *d,
                   double *dh, double *dH)
{
    /* e_iso.c:490-500 */
    const double R = 0.001;
    const int pow_d = 2;
    const double delta = npml * h;
    const double d0 = log(1. / R) * 3. / 2. * vp_mean / delta;
    for (size_t n = 0; n < npml; ++n) {
        d[n] = d0 * pow(((npml - n) * h / delta), pow_d);
        dh[n] = d0 * pow((((npml - n) * h - h / 2.) / delta), pow_d);
    }
    for (size_t n = 0; n < npml + 1; ++n)
        dH[n] = d0 * pow((((npml - n) * h + h / 2.) / delta), pow_d);
}

/* ------------------------------------------------------------------ */
/* Region tables                                                       */
/* ------------------------------------------------------------------ */
enum { R_UL, R_L, R_LL, R_U, R_LO, R_UR, R_R, R_LR, NREG };
enum { T_NONE, T_D, T_DH, T_DBIG };   /* undamped, d[], dh[], dH[] */
enum { IX_FWD, IX_REV };              /* n, Npml-1-n */
enum { C_ZERO, C_LOC, C_LAST };       /* interior coordinate 0, local, last */

typedef struct { unsigned char tab, ix; } damp;

/* Damping of the x part (index by local i) and z part (index by local j)
 * of each split pair, per region in the order UL, L, LL, U, LO, UR, R, LR. */
static const damp VX_X[NREG] = {            /* pml.c:56-75, 99-118, 134-153, 178-189, */
    {T_D, IX_FWD}, {T_D, IX_FWD}, {T_D, IX_FWD},          /* 214-225, 249-268,   */
    {T_NONE, 0}, {T_NONE, 0},                             /* 291-310, 326-345    */
    {T_D, IX_REV}, {T_D, IX_REV}, {T_D, IX_REV}};
static const damp VX_Z[NREG] = {            /* pml.c:76-94, 120-129, 155-172, 190-208, */
    {T_D, IX_FWD}, {T_NONE, 0}, {T_D, IX_REV},            /* 226-243, 269-287,   */
    {T_D, IX_FWD}, {T_D, IX_REV},                         /* 312-321, 347-364    */
    {T_D, IX_FWD}, {T_NONE, 0}, {T_D, IX_REV}};
static const damp VZ_X[NREG] = {            /* pml.c:372-392, 415-434, 449-469, 492-504, */
    {T_DH, IX_FWD}, {T_DH, IX_FWD}, {T_DH, IX_FWD},       /* 527-539, 561-581,   */
    {T_NONE, 0}, {T_NONE, 0},                             /* 605-625, 641-661    */
    {T_DBIG, IX_REV}, {T_DBIG, IX_REV}, {T_DBIG, IX_REV}};
static const damp VZ_Z[NREG] = {            /* pml.c:394-411, 436-445, 470-487, 505-522, */
    {T_DH, IX_FWD}, {T_NONE, 0}, {T_DBIG, IX_REV},        /* 540-557, 583-600,   */
    {T_DH, IX_FWD}, {T_DBIG, IX_REV},                     /* 627-636, 662-679    */
    {T_DH, IX_FWD}, {T_NONE, 0}, {T_DBIG, IX_REV}};
static const damp TN_X[NREG] = {            /* txx_x, tzz_x: pml.c:701-727, 755-780,   */
    {T_DH, IX_FWD}, {T_DH, IX_FWD}, {T_DH, IX_FWD},       /* 798-824, 854-868,   */
    {T_NONE, 0}, {T_NONE, 0},                             /* 897-911, 940-965,   */
    {T_DBIG, IX_REV}, {T_DBIG, IX_REV}, {T_DBIG, IX_REV}};/* 994-1019, 1037-1062 */
static const damp TN_Z[NREG] = {            /* txx_z, tzz_z: pml.c:728-751, 782-793,   */
    {T_D, IX_FWD}, {T_NONE, 0}, {T_D, IX_REV},            /* 826-849, 869-892,   */
    {T_D, IX_FWD}, {T_D, IX_REV},                         /* 912-935, 966-989,   */
    {T_D, IX_FWD}, {T_NONE, 0}, {T_D, IX_REV}};           /* 1021-1032, 1064-1087*/
static const damp TXZ_X[NREG] = {           /* pml.c:1095-1114, 1138-1157, 1173-1192, */
    {T_D, IX_FWD}, {T_D, IX_FWD}, {T_D, IX_FWD},          /* 1216-1227, 1251-1262, */
    {T_NONE, 0}, {T_NONE, 0},                             /* 1286-1305, 1329-1348, */
    {T_D, IX_REV}, {T_D, IX_REV}, {T_D, IX_REV}};         /* 1365-1384           */
static const damp TXZ_Z[NREG] = {           /* pml.c:1116-1133, 1159-1168, 1193-1210, */
    {T_DH, IX_FWD}, {T_NONE, 0}, {T_DBIG, IX_REV},        /* 1228-1245, 1263-1280, */
    {T_DH, IX_FWD}, {T_DBIG, IX_REV},                     /* 1307-1324, 1350-1359, */
    {T_DH, IX_FWD}, {T_NONE, 0}, {T_DBIG, IX_REV}};       /* 1385-1402           */

/* Interior material cell used by each region: (x coordinate, z coordinate)
 * e.g. b[0], b[j], b[nz-1], b[i*nz], b[i*nz+nz-1], b[(nx-1)*nz],
 * b[(nx-1)*nz+j], b[nx*nz-1] (pml.c:62, 105, 140, 187, 223, 255, 297, 332). */
static const unsigned char COEF_I[NREG] = {C_ZERO, C_ZERO, C_ZERO, C_LOC, C_LOC, C_LAST, C_LAST, C_LAST};
static const unsigned char COEF_J[NREG] = {C_ZERO, C_LOC, C_LAST, C_ZERO, C_LAST, C_ZERO, C_LOC, C_LAST};

typedef struct { size_t I0, J0, ni, nj, off; } region;

typedef struct {
    size_t nx, nz, np, NX, NZ;
    double h, dt;
    region reg[NREG];
    /* interior fields, nx*nz */
    double *vx, *vz, *txx, *tzz, *txz;
    /* split PML fields, compact region storage */
    double *vx_x, *vx_z, *vz_x, *vz_z, *txx_x, *txx_z, *tzz_x, *tzz_z, *txz_x, *txz_z;
    const double *la, *mu, *l2m, *b, *bm;
    const double *tabs[4];       /* NULL, d, dh, dH */
    size_t q1_index;             /* material cell of pml_t's lower-right corner */
} e2d_state;

static int region_of(const e2d_state *s, size_t I, size_t J)
{
    const size_t np = s->np, nx = s->nx, nz = s->nz;
    int col = I < np ? 0 : (I < np + nx ? 1 : 2);
    int row = J < np ? 0 : (J < np + nz ? 1 : 2);
    static const int map[3][3] = {{R_UL, R_L, R_LL}, {R_U, -1, R_LO}, {R_UR, R_R, R_LR}};
    return map[col][row];
}

/* total value of a field at padded (I, J); out of the grid is 0 */
static double fetch(const e2d_state *s, const double *interior,
                    const double *fx, const double *fz, long I, long J)
{
    if (I < 0 || J < 0 || (size_t)I >= s->NX || (size_t)J >= s->NZ) return 0.;
    int r = region_of(s, (size_t)I, (size_t)J);
    if (r < 0) return interior[((size_t)I - s->np) * s->nz + ((size_t)J - s->np)];
    const region *g = &s->reg[r];
    size_t c = g->off + ((size_t)I - g->I0) * g->nj + ((size_t)J - g->J0);
    return fx[c] + fz[c];
}

#define VX(I, J) fetch(s, s->vx, s->vx_x, s->vx_z, (I), (J))
#define VZ(I, J) fetch(s, s->vz, s->vz_x, s->vz_z, (I), (J))
#define TXX(I, J) fetch(s, s->txx, s->txx_x, s->txx_z, (I), (J))
#define TZZ(I, J) fetch(s, s->tzz, s->tzz_x, s->tzz_z, (I), (J))
#define TXZ(I, J) fetch(s, s->txz, s->txz_x, s->txz_z, (I), (J))

static double damping_of(const e2d_state *s, damp dm, size_t loc)
{
    size_t n = dm.ix == IX_FWD ? loc : s->np - 1 - loc;
    return s->tabs[dm.tab][n];
}

/* u = ((u*c2) + coef*(t1-t2)/h)*c1, pml.c:59-63 */
static inline double damped(double u, double dd, double coef, double t1,
                            double t2, double dt, double h)
{
    double c1 = (2. * dt) / (2. + dt * dd);
    double c2 = (2. - dt * dd) / (2. * dt);
    u *= c2;
    u += coef * (t1 - t2) / h;
    u *= c1;
    return u;
}

/* u += dt*coef*(t1-t2)/h, e.g. pml.c:123 */
static inline double undamped(double u, double coef, double t1, double t2,
                              double dt, double h)
{
    return u + dt * coef * (t1 - t2) / h;
}

static size_t coef_index(const e2d_state *s, int r, size_t i, size_t j)
{
    size_t ci = COEF_I[r] == C_ZERO ? 0 : (COEF_I[r] == C_LOC ? i : s->nx - 1);
    size_t cj = COEF_J[r] == C_ZERO ? 0 : (COEF_J[r] == C_LOC ? j : s->nz - 1);
    return ci * s->nz + cj;
}

static void split_update(const e2d_state *s, double *u, damp dm, size_t loc,
                         double coef, double t1, double t2)
{
    if (dm.tab == T_NONE) *u = undamped(*u, coef, t1, t2, s->dt, s->h);
    else *u = damped(*u, damping_of(s, dm, loc), coef, t1, t2, s->dt, s->h);
}

/* pml_v (pml.c:39-680), region by region */
static void pml_velocity(e2d_state *s)
{
    for (int r = 0; r < NREG; ++r) {
        const region *g = &s->reg[r];
        for (size_t i = 0; i < g->ni; ++i)
            for (size_t j = 0; j < g->nj; ++j) {
                const long I = (long)(g->I0 + i), J = (long)(g->J0 + j);
                const size_t c = g->off + i * g->nj + j;
                const size_t m = coef_index(s, r, i, j);
                split_update(s, &s->vx_x[c], VX_X[r], i, s->b[m], TXX(I, J), TXX(I - 1, J));
                split_update(s, &s->vx_z[c], VX_Z[r], j, s->b[m], TXZ(I, J), TXZ(I, J - 1));
                split_update(s, &s->vz_x[c], VZ_X[r], i, s->bm[m], TXZ(I + 1, J), TXZ(I, J));
                split_update(s, &s->vz_z[c], VZ_Z[r], j, s->bm[m], TZZ(I, J + 1), TZZ(I, J));
            }
    }
}

/* pml_t (pml.c:684-1403), region by region */
static void pml_stress(e2d_state *s)
{
    for (int r = 0; r < NREG; ++r) {
        const region *g = &s->reg[r];
        for (size_t i = 0; i < g->ni; ++i)
            for (size_t j = 0; j < g->nj; ++j) {
                const long I = (long)(g->I0 + i), J = (long)(g->J0 + j);
                const size_t c = g->off + i * g->nj + j;
                const size_t m = coef_index(s, r, i, j);
                const size_t mn = r == R_LR ? s->q1_index : m;   /* pml.c:1044-1084 */
                double vx1 = VX(I + 1, J), vx2 = VX(I, J);
                double vz1 = VZ(I, J), vz2 = VZ(I, J - 1);
                split_update(s, &s->txx_x[c], TN_X[r], i, s->l2m[mn], vx1, vx2);
                split_update(s, &s->tzz_x[c], TN_X[r], i, s->la[mn], vx1, vx2);
                split_update(s, &s->txx_z[c], TN_Z[r], j, s->la[mn], vz1, vz2);
                split_update(s, &s->tzz_z[c], TN_Z[r], j, s->l2m[mn], vz1, vz2);
                split_update(s, &s->txz_x[c], TXZ_X[r], i, s->mu[m], VZ(I, J), VZ(I - 1, J));
                split_update(s, &s->txz_z[c], TXZ_Z[r], j, s->mu[m], VX(I, J + 1), VX(I, J));
            }
    }
}

/* update_vx (propagate.c:2511-2550) and update_vz (2552-2588) */
static void interior_velocity(e2d_state *s, int which)
{
    const double dt = s->dt, h = s->h;
    for (size_t i = 0; i < s->nx; ++i)
        for (size_t j = 0; j < s->nz; ++j) {
            const long I = (long)(i + s->np), J = (long)(j + s->np);
            const size_t c = i * s->nz + j;
            if (which == 0) {
                double dt1 = (TXX(I, J) - TXX(I - 1, J)) / h;
                double dt2 = (TXZ(I, J) - TXZ(I, J - 1)) / h;
                s->vx[c] += dt * s->b[c] * (dt1 + dt2);
            } else {
                double dt1 = (TXZ(I + 1, J) - TXZ(I, J)) / h;
                double dt2 = (TZZ(I, J + 1) - TZZ(I, J)) / h;
                s->vz[c] += dt * s->bm[c] * (dt1 + dt2);
            }
        }
}

/* update_txxzz (propagate.c:2590-2632) and update_txz (2634-2670) */
static void interior_stress(e2d_state *s, int which)
{
    const double dt = s->dt, h = s->h;
    for (size_t i = 0; i < s->nx; ++i)
        for (size_t j = 0; j < s->nz; ++j) {
            const long I = (long)(i + s->np), J = (long)(j + s->np);
            const size_t c = i * s->nz + j;
            if (which == 0) {
                double dvx_x = dt * (VX(I + 1, J) - VX(I, J)) / h;
                double dvz_z = dt * (VZ(I, J) - VZ(I, J - 1)) / h;
                s->txx[c] += s->l2m[c] * dvx_x + s->la[c] * dvz_z;
                s->tzz[c] += s->l2m[c] * dvz_z + s->la[c] * dvx_x;
            } else {
                double dvx_z = dt * (VX(I, J + 1) - VX(I, J)) / h;
                double dvz_x = dt * (VZ(I, J) - VZ(I - 1, J)) / h;
                s->txz[c] += s->mu[c] * (dvx_z + dvz_x);
            }
        }
}

static void to_padded(const e2d_state *s, const double *interior,
                      const double *fx, const double *fz, double *out)
{
    for (size_t I = 0; I < s->NX; ++I)
        for (size_t J = 0; J < s->NZ; ++J)
            out[I * s->NZ + J] = fetch(s, interior, fx, fz, (long)I, (long)J);
}

static void split_padded(const e2d_state *s, const double *f, double *out)
{
    memset(out, 0, s->NX * s->NZ * sizeof(double));
    for (int r = 0; r < NREG; ++r) {
        const region *g = &s->reg[r];
        for (size_t i = 0; i < g->ni; ++i)
            for (size_t j = 0; j < g->nj; ++j)
                out[(g->I0 + i) * s->NZ + g->J0 + j] = f[g->off + i * g->nj + j];
    }
}

int o_e2d_run(const o_e2d_params *p, const double *vp, const double *vs,
              const double *rho, size_t nsrc, const int *src_type,
              const size_t *src_idx, const size_t *src_len,
              const double *src_fct, size_t nrec, const size_t *rec_idx,
              const int *rec_comp, size_t nsteps, double *traces,
              double *fields, double *split)
{
    if (p->nx < 2 || p->nz < 2 || p->npml < 1) return O_ERR_ARG;
    e2d_state st, *s = &st;
    const size_t nx = p->nx, nz = p->nz, np = p->npml;
    s->nx = nx; s->nz = nz; s->np = np;
    s->NX = nx + 2 * np; s->NZ = nz + 2 * np;
    s->h = p->h; s->dt = p->dt;
    const size_t n = nx * nz;
    const size_t npml_cells = 4 * np * np + 2 * np * nx + 2 * np * nz;
    const region regs[NREG] = {
        {0, 0, np, np, 0},
        {0, np, np, nz, np * np},
        {0, np + nz, np, np, np * np + np * nz},
        {np, 0, nx, np, 2 * np * np + np * nz},
        {np, np + nz, nx, np, 2 * np * np + np * nz + np * nx},
        {np + nx, 0, np, np, 2 * np * np + np * nz + 2 * np * nx},
        {np + nx, np, np, nz, 3 * np * np + np * nz + 2 * np * nx},
        {np + nx, np + nz, np, np, 3 * np * np + 2 * np * nz + 2 * np * nx}};
    memcpy(s->reg, regs, sizeof(regs));

    double *mat = malloc(5 * n * sizeof(double));
    double *fld = calloc(5 * n + 10 * npml_cells, sizeof(double));
    double *tab = calloc(3 * np + 1, sizeof(double));
    if (!mat || !fld || !tab) { free(mat); free(fld); free(tab); return O_ERR_NOMEM; }
    double *la = mat, *mu = mat + n, *l2m = mat + 2 * n, *b = mat + 3 * n, *bm = mat + 4 * n;
    o_e2d_material(nx, nz, vp, vs, rho, la, mu, l2m, b, bm);
    s->la = la; s->mu = mu; s->l2m = l2m; s->b = b; s->bm = bm;

    /* e_iso.c:423-427, 476-479: CFL from the maximum Vp */
    double vpmax = 0.0;
    for (size_t c = 0; c < n; ++c) vpmax = vp[c] > vpmax ? vp[c] : vpmax;
    if (p->dt > p->h / (sqrt(2.0) * vpmax)) {
        free(mat); free(fld); free(tab);
        return O_ERR_CFL;
    }
    /* e_iso.c:486-488: damping from the MEAN Vp */
    double vmean = 0.0;
    for (size_t c = 0; c < n; ++c) vmean += vp[c];
    vmean /= (double)nx * nz;
    double *d = tab, *dh = tab + np, *dH = tab + 2 * np;
    o_e2d_damping(np, p->h, vmean, d, dh, dH);
    if (p->damping_scale != 1.0)
        for (size_t c = 0; c < 3 * np + 1; ++c) tab[c] *= p->damping_scale;
    s->tabs[0] = NULL; s->tabs[1] = d; s->tabs[2] = dh; s->tabs[3] = dH;
    /* pml.c:1044-1084 reads l2m/la[nx*nx-1]; where that is past the array
     * (nx > nz) the reference is undefined and the true corner is used. */
    s->q1_index = (p->q1_compat && nx * nx <= n) ? nx * nx - 1 : n - 1;

    double *f = fld;
    s->vx = f; s->vz = f + n; s->txx = f + 2 * n; s->tzz = f + 3 * n; s->txz = f + 4 * n;
    f += 5 * n;
    double **sp[10] = {&s->vx_x, &s->vx_z, &s->vz_x, &s->vz_z, &s->txx_x,
                       &s->txx_z, &s->tzz_x, &s->tzz_z, &s->txz_x, &s->txz_z};
    for (int k = 0; k < 10; ++k) *sp[k] = f + k * npml_cells;

    const int nth = p->nthreads > 0 ? p->nthreads : 1;
    (void)nth;
    for (size_t it = 0; it < nsteps; ++it) {
        /* e_iso.c:595-615: vx, vz and pml_v as three sections */
#pragma omp parallel sections num_threads(nth)
        {
#pragma omp section
            interior_velocity(s, 0);
#pragma omp section
            interior_velocity(s, 1);
#pragma omp section
            pml_velocity(s);
        }
        /* e_iso.c:616-634 */
#pragma omp parallel sections num_threads(nth)
        {
#pragma omp section
            interior_stress(s, 0);
#pragma omp section
            interior_stress(s, 1);
#pragma omp section
            pml_stress(s);
        }
        /* add_src, src.c:170-184 */
        size_t off = 0;
        for (size_t k = 0; k < nsrc; ++k) {
            if (it < src_len[k]) {
                double inc = p->dt * src_fct[off + it];
                int ty = src_type[k];
                if (ty == O_SX || ty == O_BULK) s->txx[src_idx[k]] += inc;
                if (ty == O_SZ || ty == O_BULK) s->tzz[src_idx[k]] += inc;
                if (ty == O_SXZ) s->txz[src_idx[k]] += inc;
            }
            off += src_len[k];
        }
        if (traces) {
            for (size_t r = 0; r < nrec; ++r) {
                const double *fsel = rec_comp[r] == O_VX ? s->vx
                                   : rec_comp[r] == O_VZ ? s->vz
                                   : rec_comp[r] == O_TXX ? s->txx
                                   : rec_comp[r] == O_TZZ ? s->tzz : s->txz;
                traces[it * nrec + r] = fsel[rec_idx[r]];
            }
        }
    }
    if (fields) {
        const size_t N = s->NX * s->NZ;
        to_padded(s, s->vx, s->vx_x, s->vx_z, fields);
        to_padded(s, s->vz, s->vz_x, s->vz_z, fields + N);
        to_padded(s, s->txx, s->txx_x, s->txx_z, fields + 2 * N);
        to_padded(s, s->tzz, s->tzz_x, s->tzz_z, fields + 3 * N);
        to_padded(s, s->txz, s->txz_x, s->txz_z, fields + 4 * N);
    }
    if (split) {
        const size_t N = s->NX * s->NZ;
        for (int k = 0; k < 10; ++k) split_padded(s, *sp[k], split + k * N);
    }
    free(mat); free(fld); free(tab);
    return O_OK;
}
