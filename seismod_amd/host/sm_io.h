/*
 * sm_io.h -- input and output of the drop-in e_iso and a_iso_3d programs.
 *
 * The file formats, keywords, defaults and messages follow seismod's
 * io_utils.c; each function names the routine it stands in for.
 */
#ifndef SM_IO_H
#define SM_IO_H

#include <stddef.h>
#include <stdio.h>

#define SM_NAME_LEN 80

enum { REC_TRACE = 0, REC_SNAPSHOT = 1 };

/* run parameters of the .par file (struct inputParams, structs.h:80-105) */
typedef struct {
    double dt;          /* s */
    double duration;    /* s */
    char modelfile[SM_NAME_LEN], sourcefile[SM_NAME_LEN], outputfile[SM_NAME_LEN];
    char basename[SM_NAME_LEN];
    int iwipe;          /* abs */
    int plot_strips;    /* plstr */
    int segy;
    int shotpt_no;
    int simulate_cmp;
    /* absorbing layers (struct abs_params) */
    size_t np;
    double vmax, pml_order, Rc, kappa_max;
    int alpha_max;
    int mode;           /* azimuthal_mode (e_cyl_iso) */
    /* keywords of the poroviscoelastic programs (io_utils.c:257-280) */
    int check_model;    /* phase velocities and attenuation: 1 print, 2 also store */
    int chkpt_inc;      /* checkpoint increment (0: none) */
    int pml_median;     /* median smoothing of the material in the layers */
    int save_energy;    /* saveEnergy: kinetic energy of a region per step */
    double roi[4];      /* EnergyROI xmin/xmax/zmin/zmax (m) */
} sm_params;

typedef struct {
    size_t nx, ny, nz;      /* ny = 1 for 2D */
    double dx, dy, dz;
    double x0, y0, z0;
} sm_grid;

typedef struct {
    int type;               /* enum sourceType numbering */
    double A, f, x, y, z;
    size_t idx;
} sm_source;

typedef struct {
    size_t nsrc, ntemplate;
    sm_source *s;
} sm_sources;

typedef struct {
    int comp;               /* enum component numbering */
    int type;               /* REC_TRACE / REC_SNAPSHOT */
    double x, y, z, t0, dt;
    size_t idx;
    FILE *fid;
    float *data;            /* SEG-Y samples */
    size_t nsamples;
} sm_record;

typedef struct {
    size_t nrec;
    sm_record *r;
    char basename[SM_NAME_LEN];
} sm_records;

extern int sm_verbose;

/* set_defaults / set_defaults_3d (io_utils.c:63-135) */
void sm_set_defaults(sm_params *p, int three_d);
/* process_args (io_utils.c:156-294): exits like the reference on -h,
 * missing -p, unreadable .par; -c (checkpoint restart) is not supported
 * by these programs and is rejected */
void sm_process_args(int argc, char *const argv[], sm_params *p, const char *usage_text);
/* the same for programs with checkpoints: -c names a checkpoint file, which
 * is copied to ckfile and then 1 is returned without reading a parameter
 * file (-p is ignored, as in the reference); 0 otherwise */
int sm_process_args_ck(int argc, char *const argv[], sm_params *p, const char *usage_text, char *ckfile,
                       size_t ckfile_len);
/* read_grid_params / read_grid3d_params (io_utils.c:296-328, 384-401) */
void sm_read_grid2d(const char *fn, sm_grid *g);
void sm_read_grid3d(const char *fn, sm_grid *g);
/* read_model_e (io_utils.c:892-947): nx*nz arrays, with the 1-D shortcut */
void sm_read_model_e(const char *fn, const sm_grid *g, double *vp, double *vs, double *rho);
/* read_model_a (io_utils.c:1029-1153): padded arrays, edges replicated */
void sm_read_model_a(const char *fn, const sm_grid *g, size_t np, double *vp, double *rho);
/* model files in the format the readers take, values with `digits`
 * significant digits (clamped to 1..17; 17: read back bit for bit); 0 on
 * success */
int sm_write_model_a(const char *fn, size_t nx, size_t ny, size_t nz, double dx, double dy,
                     double dz, double x0, double y0, double z0, const double *vp,
                     const double *rho, int digits);
int sm_write_model_e(const char *fn, size_t nx, size_t nz, double dx, double dz, double x0,
                     double z0, const double *vp, const double *vs, const double *rho, int digits);
/* read_source / read_source_3d (io_utils.c:1432-1589) */
void sm_read_source2d(const char *fn, const sm_grid *g, sm_sources *src);
/* read_source on a grid padded by np nodes on every side (ve_vti_sh_pml) */
void sm_read_source2d_np(const char *fn, const sm_grid *g, size_t np, sm_sources *src);
void sm_read_source3d(const char *fn, const sm_grid *g, size_t np, sm_sources *src);
/* read_output / read_output_3d (io_utils.c:1592-1741) */
void sm_read_output2d(const char *fn, const sm_grid *g, sm_records *out);
void sm_read_output2d_np(const char *fn, const sm_grid *g, size_t np, sm_records *out);
/* read_grid_params_ve_sh / read_model_ve_sh (io_utils.c:330-354, 1316-1429):
 * padded (nx + 2np) x (nz + 2np) arrays, mu in GPa */
void sm_read_grid_ve_sh(const char *fn, sm_grid *g);
void sm_read_model_ve_sh(const char *fn, const sm_grid *g, size_t np, double *mu, double *rho,
                         double *gamma, double *Q, double *fr);
/* read_grid_params_ve / read_model_ve (io_utils.c:356-382, 1155-1314):
 * the grid and the number L of quasi-dilatational relaxation mechanisms;
 * padded (nx + 2np) x (nz + 2np) arrays, K and mu in GPa, Q1 and f1 holding
 * L arrays one after the other */
void sm_read_grid_ve(const char *fn, sm_grid *g, int *L);
void sm_read_model_ve(const char *fn, const sm_grid *g, size_t np, int L, double *K, double *mu,
                      double *rho, double *eps, double *delta, double *Q1, double *f1, double *Q2,
                      double *f2);
/* read_grid_params / read_model (io_utils.c:296-328, 403-625) for the
 * poroviscoelastic programs: type "isotropic" (nx x nz nodes, or one z
 * column, values copied into the absorbing layers from the nearest model
 * node) or "isotropic_pml" (the padded grid in the file: nx and nz include
 * the 2 np layer nodes, which read_grid_params subtracts); 12 values per
 * node "K_m K_s K_f phi mu rho_s rho_f T eta kappa Q f0", x outer and z
 * inner, into padded (nx + 2np) x (nz + 2np) arrays.  *with_pml tells
 * which type the file holds. */
void sm_read_grid_pve(const char *fn, size_t np, sm_grid *g, int *with_pml);
void sm_read_model_pve(const char *fn, const sm_grid *g, size_t np, int with_pml, double *const vals[12]);
/* read_grid_params / read_modelVTI (io_utils.c:296-328, 627-888) for
 * pve_vti_pml: type "anisotropic_vti" (nx x nz nodes or one z column) or
 * "anisotropic__vti_pml" (the padded grid in the file; nx and nz still
 * count interior nodes); 16 values per node "K_m K_s K_f phi mu rho_s rho_f
 * T3 eta kappa3 Q f0 epsilon delta kappa3/kappa1 T3/T1", returned in vals
 * as K_m K_s K_f phi mu rho_s rho_f T1 T3 eta kappa1 kappa3 Q f0 epsilon
 * delta (padded arrays) */
void sm_read_grid_pve_vti(const char *fn, sm_grid *g, int *with_pml);
void sm_read_model_pve_vti(const char *fn, const sm_grid *g, size_t np, int with_pml, double *const vals[16]);
/* Phase velocity (m/s) and attenuation (17.372 pi Im V / Re V) of the slow
 * and fast P waves and the S wave at frequency f (Hz) for n nodes of the
 * 12 model values as sm_read_model_pve gives them, from the complex moduli
 * and densities of the poroviscoelastic model (pve_iso_pml.c:864-904):
 * vel[0..2] and att[0..2] receive n values each (slow P, fast P, S). */
void sm_pve_dispersion(size_t n, double *const vals[12], double f, double *const vel[3], double *const att[3]);
/* pve_vti_pml's check_model (pve_vti_pml.c:931-1100): phase velocity (m/s)
 * and attenuation (-omega Im 1/V) of the fast and slow qP and the qS waves
 * along x and along z at frequency f, for the padded nx2 x nz2 grid of the
 * 16 VTI model values (sm_read_model_pve_vti's order); vel and att receive
 * qP1_x, qP2_x, qS_x, qP1_z, qP2_z, qS_z */
void sm_pve_dispersion_vti(size_t nx2, size_t nz2, double *const vals[16], double f, double *const vel[6],
                           double *const att[6]);
void sm_read_output3d(const char *fn, const sm_grid *g, size_t np, sm_records *out);
/* check_dt_trc (io_utils.c:3561-3574) */
void sm_check_dt_trc(const sm_records *out, double dt);
/* write_trace (io_utils.c:1744-1768): one "%le    %le" line, t and 1e3*v */
void sm_write_trace(sm_records *out, size_t nr, double t, double value);
void sm_close_traces(sm_records *out);
/* write_snapshot_nc / write_snapshot3d_nc (io_utils.c:1822-2180): data is
 * an nxp x nzp i-major slice (nxp = nx+2np, nzp = nz+2np) */
int sm_write_snapshot_nc(const double *data, double t, size_t nx, size_t nz, size_t np,
                         double x0, double z0, double dx, double dz, const char *basename,
                         size_t nr, int plstr);
/* write_field_nc (io_utils.c:2183-2352): a padded nx2 x nz2 field written
 * unscaled to <basename>_<field>.nc in the current directory, with the
 * given units */
int sm_write_field_nc(const double *data, const char *field, const char *units, size_t nx, size_t nz,
                      size_t np, double x0, double z0, double dx, double dz, const char *basename,
                      int plstr);
/* the same for e_cyl_iso's cylindrical grid: data is (nx+np) x (nz+2np) */
int sm_write_snapshot_cyl_nc(const double *data, double t, size_t nx, size_t nz, size_t np,
                             double x0, double z0, double dx, double dz, const char *basename,
                             size_t nr, int plstr);
/* e_cyl_iso inputs on the (nx+np) x (nz+2np) grid, index i*nz2 + j:
 * read_model_e_cyl (io_utils.c:949-1027), read_source and read_output
 * with the cylindrical index rules (io_utils.c:1432-1538, 1592-1682) */
void sm_read_model_e_cyl(const char *fn, const sm_grid *g, size_t np, double *vp, double *vs,
                         double *rho);
void sm_read_source_cyl(const char *fn, const sm_grid *g, size_t np, sm_sources *src);
void sm_read_output_cyl(const char *fn, const sm_grid *g, size_t np, sm_records *out);
/* save_segy (io_utils.c:3269-3559) */
int sm_save_segy(const sm_params *p, const sm_records *out, const sm_sources *src);
void sm_free_sources(sm_sources *s);
void sm_free_records(sm_records *r);

#endif
