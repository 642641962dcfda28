/*
 * oracle.h -- FP64 CPU restatement of seismod's two FDTD time loops.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP library in
 * seismod_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.  The product never links it.
 *
 * It restates, in double precision and in the reference's own storage
 * layouts, the arithmetic of
 *   - e_iso.c:423-500, 565-670, propagate.c:2511-2670, pml.c:39-1403,
 *     src.c:43-65 and 170-184 (2D elastic, split PML);
 *   - a_iso_3d.c:368-417, 494-693 (3D acoustic, CPML);
 *   - the record predicate of e_iso.c:643-658 / a_iso_3d.c:700-718 and the
 *     step count of e_iso.c:575-576.
 * Every entry point returns 0 on success and a non-zero code on a
 * condition where the reference aborts or exits (CFL, bad sizes).
 *
 * Pinning status of each function is stated beside its declaration and in
 * the tests that check it (tests/test_oracle_*.py).
 */
#ifndef SEISMOD_ORACLE_H
#define SEISMOD_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define O_OK 0
#define O_ERR_ARG 1
#define O_ERR_CFL 3
#define O_ERR_NOMEM 4

/* Source and record enums use the numbering of structs.h:44-45. */
enum o_srctype { O_SX = 0, O_SY, O_SZ, O_SXY, O_SXZ, O_SYZ, O_BULK, O_SF };
enum o_comp { O_VX = 0, O_VY = 1, O_VZ = 2, O_TXX = 5, O_TZZ = 6, O_TXZ = 7, O_P = 8 };

/* Ricker length ceil(2/(f dt)) (src.c:50-51).  Pinned: closed form. */
size_t o_ricker_length(double f, double dt);

/* compute_src_fct (src.c:43-65), including its normalisation through
 * source 0 (every source's loop rescales fct of source 0).  fct holds the
 * concatenated functions, source n starting at sum(length[0..n-1]).
 * Reading fct_0 beyond its length (a later source with a longer wavelet)
 * is undefined in the reference; here such entries are skipped.
 * Pinned: single source against the textbook Ricker; multi-source
 * against the closed form of the un-normalised later sources. */
void o_compute_src_fct(size_t n, const double *A, const double *f, double dt,
                       const size_t *length, double *fct);

/* nsteps rule of e_iso.c:575-576 / a_iso_3d.c:578-579.  Pinned: bracket
 * property (nsteps-1)*dt < duration <= nsteps*dt. */
size_t o_nsteps(double duration, double dt);

/* Record predicate of e_iso.c:643-658 for step it (t = dt*(it+1)). */
int o_record_due(size_t it, double dt, double t0, double rdt);

/* ------------------------------------------------------------------ */
/* 3D acoustic CPML (a_iso_3d)                                         */
/* ------------------------------------------------------------------ */
typedef struct {
    size_t nx, ny, nz;     /* interior nodes */
    size_t np;             /* CPML layers */
    double dx, dy, dz;
    double dt;
    double vmax, pml_order, Rc, kappa_max;
    int alpha_max;         /* 0: alpha = 0; otherwise alpha_max = pi*f0 */
    double f0;
    int nthreads;          /* OpenMP threads for the velocity sections */
    int pressure_threads;  /* 1 = single section as in a_iso_3d.c:655-684 */
    int split_threads;     /* > 0: instead, every sweep split over x planes on
                              this many threads (same values, all cores) */
} o_a3d_params;

/* CPML tables of a_iso_3d.c:500-576 for one axis of n2 nodes.  node
 * tables have n2 entries, half tables n2-1.  Pinned: closed form. */
void o_cpml_tables(size_t n2, size_t np, double d, double dt, double vmax,
                   double pml_order, double Rc, double kappa_max,
                   double alpha_max, double *a, double *b, double *k,
                   double *a_h, double *b_h, double *k_h);

/* Whole a_iso_3d run from the padded model (as read_model_a leaves it).
 *   traces: [nsteps][nrec], field of record r after step it (NULL: none)
 *   fields: P, vx, vy, vz concatenated, padded layout (NULL: none)
 * Returns O_ERR_CFL where a_iso_3d.c:411-417 aborts. */
int o_a3d_run(const o_a3d_params *p, const double *vp, const double *rho,
              size_t nsrc, const size_t *src_idx, const size_t *src_len,
              const double *src_fct, size_t nrec, const size_t *rec_idx,
              const int *rec_comp, size_t nsteps, double *traces,
              double *fields);

/* ------------------------------------------------------------------ */
/* 2D elastic split PML (e_iso)                                        */
/* ------------------------------------------------------------------ */
typedef struct {
    size_t nx, nz;   /* interior nodes */
    size_t npml;     /* PML layers (must be >= 1, see e_iso.c:346) */
    double h, dt;
    int q1_compat;   /* use l2m/la[nx*nx-1] in the lower-right corner of
                        pml_t (pml.c:1044-1084) when nx*nx <= nx*nz */
    int nthreads;    /* OpenMP threads for the three sections */
    double damping_scale; /* 1.0 for the reference; tests use 0.0 */
} o_e2d_params;

/* Damping tables d[Np], dh[Np], dH[Np+1] of e_iso.c:482-500 from the
 * mean Vp.  Pinned: closed form. */
void o_e2d_damping(size_t npml, double h, double vp_mean, double *d,
                   double *dh, double *dH);

/* Staggered material of e_iso.c:423-462: la, mu, l2m, b, bm (nx*nz). */
void o_e2d_material(size_t nx, size_t nz, const double *vp, const double *vs,
                    const double *rho, double *la, double *mu, double *l2m,
                    double *b, double *bm);

/* Whole e_iso run.
 *   src_type uses enum o_srctype; src_idx are interior indices i*nz+j.
 *   traces: [nsteps][nrec] (VX, VZ, TXX, TZZ, TXZ on the interior arrays)
 *   fields: vx, vz, txx, tzz, txz totals on the padded (nx+2Np)x(nz+2Np)
 *           grid (PML cells hold x+z of the split pair), concatenated.
 *   split:  the ten split arrays, each padded, zero in the interior
 *           (order vx_x vx_z vz_x vz_z txx_x txx_z tzz_x tzz_z txz_x txz_z)
 * Returns O_ERR_CFL where e_iso.c:476-479 aborts. */
int o_e2d_run(const o_e2d_params *p, const double *vp, const double *vs,
              const double *rho, size_t nsrc, const int *src_type,
              const size_t *src_idx, const size_t *src_len,
              const double *src_fct, size_t nrec, const size_t *rec_idx,
              const int *rec_comp, size_t nsteps, double *traces,
              double *fields, double *split);

/* ------------------------------------------------------------------ */
/* 2.5D elastic, cylindrical coordinates, CPML (e_cyl_iso)             */
/* ------------------------------------------------------------------ */
/* source types and record components of e_cyl_iso (structs.h:44-45) */
enum { O_SRC_BULK = 6, O_SRC_FR = 9, O_SRC_FT = 10, O_SRC_FZ = 11, O_SRC_KURKJIAN = 12 };
enum { O_DIV = 11, O_CURL = 12, O_VR = 13, O_VT = 14, O_TRR = 15, O_TRT = 16, O_TRZ = 17,
       O_TTT = 18, O_TTZ = 19 };

typedef struct {
    size_t nx, nz;       /* interior nodes along r and z */
    size_t np;           /* CPML layers (0: abs off); nx2 = nx+np, nz2 = nz+2np */
    double dx, dz, dt;
    int n;               /* azimuthal mode (azimuthal_mode, >= 0) */
    double vmax, pml_order, Rc, kappa_max;
    double alpha_max;    /* pi f0, or 0 when abs_params.alpha_max is 0 */
} OCylParams;

/* Staggered material of e_cyl_iso.c:342-415 on the padded nx2 x nz2 grid
 * (lambda, lambda+2mu at cell centres, mu at nodes by 4-point harmonic
 * means, mu and 1/rho averaged between r and z neighbours).  Returns 0, or
 * 1 + the first cell holding a non-finite value (the reference aborts).
 * Pinned: closed forms (tests/test_oracle_cyl.py). */
int o_cyl_material(size_t nx2, size_t nz2, const double *vp, const double *vs, const double *rho,
                   double *lij, double *l2mij, double *mu, double *mui, double *muj, double *bi,
                   double *bj, double *bij, double *vpmax);

/* CPML tables of pml.c:2107-2164 (see oracle_cyl.c for the order in t:
 * 9 tables of np then 6 of np+1).  Pinned: closed form. */
void o_cyl_cpml(size_t np, double dx, double dz, double dt, double vmax, double order, double Rc,
                double kappa_max, double alpha_max, double *t);

/* Kurkjian source function of src.c:202-229.  Pinned: closed form. */
void o_kurkjian_fct(double lij_src, double dx, double dz, double x, int n, double A, double f,
                    double dt, size_t length, double *fct);

/* Whole e_cyl_iso run (e_cyl_iso.c:601-866).
 *   vp, vs, rho: padded nx2 x nz2 model as read_model_e_cyl leaves it;
 *   type: O_SRC_*; force sources (Fr, Ftheta, Fz, Bulk) after the velocity
 *   update, a Kurkjian source as entry 0 after the stress update;
 *   rcomp: O_VR .. O_TTZ, O_VZ, O_TZZ, O_DIV, O_CURL;
 *   fields: vr, vt, vz, trr, ttt, tzz, trz, trt, ttz (nx2*nz2 each).
 * O_ERR_CFL where e_cyl_iso.c:454-458 aborts, O_ERR_ARG for a material
 * value that is not finite (417-452). */
int o_cyl_run(const OCylParams *p, const double *vp, const double *vs, const double *rho,
              size_t nsrc, const int *type, const size_t *idx, const size_t *len, const double *fct,
              size_t nrec, const size_t *ridx, const int *rcomp, size_t nsteps, double *traces,
              double *fields);

#ifdef __cplusplus
}
#endif
#endif
