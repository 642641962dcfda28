/*
 * pve_driver.c -- the programs pve_iso_pml and pve_vti_pml: poroviscoelastic
 * wave propagation in 2D isotropic or VTI media (Biot solid and fluid
 * motion after Carcione and Helle 1999, one relaxation mechanism for the
 * fluid pressure, pseudo-spectral derivatives, CPML, fourth-order
 * Runge-Kutta with the viscous coupling stepped analytically), running the
 * time loop on an MI355X through libseismod_hip.
 *
 * Drop-in for seismod's pve_iso_pml (pve_iso_pml.c:441-1440) and
 * pve_vti_pml (pve_vti_pml.c:441-1400): same command line, .par keywords,
 * model files ("isotropic"/"isotropic_pml", 12 values per node, or
 * "anisotropic_vti"/"anisotropic__vti_pml", 16), source (Sx, Sz, Sxz, Bulk,
 * Sf, Bulk_s; 1- or 9-point template) and record (Vx, Vz, qx, qz, txx, tzz,
 * txz, p, and for the isotropic program div and curl) files, and the same
 * outputs: trc/, ssh/, segy/, the kinetic energy of a region in
 * E/<base>_E.dat (saveEnergy, EnergyROI) and the phase velocity and
 * attenuation check (check_model: 1 prints the extremes, 2 also writes
 * <base>_vel_slowP.nc ... <base>_att_S.nc, or for the VTI program
 * <base>_vel_qP1_x.nc ... <base>_att_qS_z.nc); pml_median 1 smooths the
 * isotropic medium's layers (check_model then still reports the unsmoothed
 * model).  chkpt_inc > 0 saves chkpt/<base>_<it>.dat after every
 * chkpt_inc-th iteration and -c <file> continues a run from one; the
 * checkpoint holds the parameters and the solver state (fields, CPML memory
 * variables, the traces and energy so far), not the inputs, so the model,
 * source and output files it names must still be there at the restart.
 *
 * Set-up and update arithmetic live in the library (sm_pve_*); this driver
 * reads the inputs, drives the time loop and writes the records.
 *
 * Environment: SEISMOD_STORAGE=f32 runs the FP32 kernels; the default is
 * FP64 (the reference's precision).  SEISMOD_DEVICE selects the GPU.
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "pve_driver.h"
#include "seismod_hip.h"
#include "sm_io.h"

static const char *usage_iso =
    "Format of parameter file:\n\n"
    "One line per parameter, first word is the value of the parameter,\n"
    "next is a keyword comprise between a # and a comma, and a comment\n"
    "can be added after the comma, i.e we have on each line\n\n"
    "value     # keyword, optional comment\n\n"
    "Keywords: model, source, output, basename, time (s), dt (ms), segy, shotpoint,\n"
    "simulateCMP, abs, nl, plstr, kappa_max, alpha_max, Rc, check_model (0, 1 display,\n"
    "2 store), saveEnergy, EnergyROI (xmin/xmax/zmin/zmax)\n\n"
    "Model file: nx nz / dx dz / x0 z0 / isotropic (or isotropic_pml: the absorbing\n"
    "layers included) / then one line per node, x outer and z inner:\n"
    "K_m[GPa] K_s[GPa] K_f[GPa] phi mu[GPa] rho_s[kg/m3] rho_f[kg/m3] T eta[cP]\n"
    "kappa[mD] Q f0[Hz]  (one column only: laterally uniform)\n\n"
    "Source file: nsrc / template points (1 or 9) / then per source: type\n"
    "(Sx, Sz, Sxz, Bulk, Sf, Bulk_s), strength, frequency [Hz], x z [m]\n\n"
    "Record file: nrec / then per record: component (Vx, Vz, qx, qz, txx, tzz, txz,\n"
    "p, div, curl), Trace x z t0[s] dt[ms] or Snapshot t0[s] dt[ms]\n";

static void check(int st, const char *where)
{
    if (st != SM_OK) {
        fprintf(stderr, "Error in %s: %s (%s)\n", where, sm_strerror(st), sm_last_error());
        exit(1);
    }
}

static const char *usage_vti =
    "Format of parameter file:\n\n"
    "One line per parameter, first word is the value of the parameter,\n"
    "next is a keyword comprise between a # and a comma, and a comment\n"
    "can be added after the comma, i.e we have on each line\n\n"
    "value     # keyword, optional comment\n\n"
    "Keywords: model, source, output, basename, time (s), dt (ms), segy, shotpoint,\n"
    "simulateCMP, abs, nl, plstr, kappa_max, alpha_max, Rc, check_model (0, 1 display,\n"
    "2 store), saveEnergy, EnergyROI (xmin/xmax/zmin/zmax)\n\n"
    "Model file: nx nz / dx dz / x0 z0 / anisotropic_vti (or anisotropic__vti_pml: the\n"
    "absorbing layers included, nx nz still the interior nodes) / then one line per\n"
    "node, x outer and z inner: K_m[GPa] K_s[GPa] K_f[GPa] phi mu[GPa] rho_s[kg/m3]\n"
    "rho_f[kg/m3] T3 eta[cP] kappa3[mD] Q f0[Hz] epsilon delta kappa3/kappa1 T3/T1\n"
    "(one column only: laterally uniform)\n\n"
    "Source file: nsrc / template points (1 or 9) / then per source: type\n"
    "(Sx, Sz, Sxz, Bulk, Sf, Bulk_s), strength, frequency [Hz], x z [m]\n\n"
    "Record file: nrec / then per record: component (Vx, Vz, qx, qz, txx, tzz, txz,\n"
    "p), Trace x z t0[s] dt[ms] or Snapshot t0[s] dt[ms]\n";

/* components the program records: pve_vti_pml has no div and curl
 * (pve_vti_pml.c:1290-1318) */
static int recorded(int c, int vti)
{
    return c == SM_VX || c == SM_VZ || c == SM_QX || c == SM_QZ || c == SM_TXX || c == SM_TZZ || c == SM_TXZ ||
           c == SM_P || (!vti && (c == SM_DIV || c == SM_CURL));
}

/* the 12 isotropic model values */
enum { KM, KS, KF, PHI, MU, RHOS, RHOF, TORT, ETA, KAPPA, QQ, FQ };
/* the 16 VTI model values (sm_read_model_pve_vti's order) */
enum { V_KM, V_KS, V_KF, V_PHI, V_MU, V_RHOS, V_RHOF, V_T1, V_T3, V_ETA, V_K1, V_K3, V_Q, V_FQ, V_EPS, V_DELTA,
       V_N };

/* Extremes of the phase velocities and attenuation of the slow and fast P
 * waves and the S wave at 1 Hz, the source frequency and 1/dt, printed when
 * verbose; check_model 2 also writes the values at the source frequency
 * (pve_iso_pml.c:846-1006 and 1164-1185). */
static void check_model(const sm_params *p, const sm_grid *g, size_t np, double *const v[12], double f0)
{
    const size_t n = (g->nx + 2 * np) * (g->nz + 2 * np);
    double ext[6][6];   /* slow P, fast P, S velocity, then attenuation: max at 3 frequencies, then min */
    for (int k = 0; k < 6; ++k)
        for (int q = 0; q < 6; ++q) ext[k][q] = q < 3 ? 0.0 : 1.e9;
    double *w[6];
    for (int k = 0; k < 6; ++k) w[k] = malloc(n * sizeof(double));
    const double freq[3] = {1.0, f0, 1. / p->dt};
    /* the source frequency last, so that w holds its values for the files */
    for (int q = 0; q < 3; ++q) {
        const int fq = q == 2 ? 1 : q == 1 ? 2 : 0;
        sm_pve_dispersion(n, v, freq[fq], w, w + 3);
        for (int k = 0; k < 6; ++k)
            for (size_t c = 0; c < n; ++c) {
                if (w[k][c] > ext[k][fq]) ext[k][fq] = w[k][c];
                if (w[k][c] < ext[k][fq + 3]) ext[k][fq + 3] = w[k][c];
            }
    }
    if (p->check_model > 1) {
        if (sm_verbose > 1) {
            fprintf(stdout, "  Saving phase velocity and attenuation @ src frequency ...");
            fflush(stdout);
        }
        const char *names[6] = {"vel_slowP", "vel_fastP", "vel_S", "att_slowP", "att_fastP", "att_S"};
        for (int k = 0; k < 6; ++k)
            sm_write_field_nc(w[k], names[k], k < 3 ? "m/s" : "1/m", g->nx, g->nz, np, g->x0, g->z0, g->dx,
                              g->dz, p->basename, 1);
    }
    for (int k = 0; k < 6; ++k) free(w[k]);
    if (sm_verbose) {
        const char *label[3] = {"   Low frequency (1 Hz)    - ", "   Src frequency (%lg Hz)   - ",
                                "  High frequency (%lg Hz) - "};
        const char *wave[3] = {"slow P", "fast P", "     S"};
        for (int a = 0; a < 2; ++a) {
            fprintf(stdout, a == 0 ? "Max/min phase velocities:\n" : "Max/min attenuation coefficient:\n");
            for (int q = 0; q < 3; ++q)
                for (int k = 0; k < 3; ++k) {
                    if (k == 0) fprintf(stdout, label[q], freq[q]);
                    else fprintf(stdout, "                             ");
                    fprintf(stdout, "%s : %lf    %lf\n", wave[k], ext[3 * a + k][q], ext[3 * a + k][q + 3]);
                }
        }
        fflush(stdout);
    }
}

/* pve_vti_pml's check_model (pve_vti_pml.c:931-1100): a warning when the
 * frame is stiffer than the solid (alpha1 or alpha3 < 0), then the extremes
 * of the qP1, qP2 and qS phase velocities and attenuation along x and z at
 * the source frequency; check_model 2 also writes them */
static void check_model_vti(const sm_params *p, const sm_grid *g, size_t np, double *const v[V_N], double f0)
{
    const size_t nx2 = g->nx + 2 * np, nz2 = g->nz + 2 * np, n = nx2 * nz2;
    for (size_t c = 0; c < n; ++c) {
        const double K_s = v[V_KS][c] * 1000., mu = v[V_MU][c] * 1000., c33 = v[V_KM][c] * 1000. + 4. / 3. * mu;
        const double c11 = 2. * v[V_EPS][c] * c33 + c33, c12 = c11 - 2. * mu;
        const double c13 = sqrt(2. * v[V_DELTA][c] * c33 * (c33 - mu) + (c33 - mu) * (c33 - mu)) - mu;
        if (1. - (c11 + c12 + c13) / (3. * K_s) < 0.0 || 1. - (2. * c13 + c33) / (3. * K_s) < 0.0) {
            fprintf(stdout, "\n***\nWarning: bulk modulus of drained matrix larger than modulus of the solid\n***\n");
            break;
        }
    }
    if (sm_verbose) {
        fprintf(stdout, "done.\nComputing phase velocity and attenuation ... ");
        fflush(stdout);
    }
    double *vel[6], *att[6];
    for (int k = 0; k < 6; ++k) { vel[k] = malloc(n * sizeof(double)); att[k] = malloc(n * sizeof(double)); }
    sm_pve_dispersion_vti(nx2, nz2, v, f0, vel, att);
    double ext[12][2];   /* vel then att of qP1_x .. qS_z: max, min */
    for (int k = 0; k < 12; ++k) {
        const double *a = k < 6 ? vel[k] : att[k - 6];
        ext[k][0] = -DBL_MAX; ext[k][1] = DBL_MAX;
        for (size_t c = 0; c < n; ++c) {
            if (a[c] > ext[k][0]) ext[k][0] = a[c];
            if (a[c] < ext[k][1]) ext[k][1] = a[c];
        }
    }
    if (sm_verbose) {
        fprintf(stdout, "done.\n");
        for (int ax = 0; ax < 2; ++ax)
            for (int w = 0; w < 2; ++w) {
                const int b = 6 * w + 3 * ax;   /* qP1, qP2, qS of this axis */
                fprintf(stdout, w == 0 ? "Max/min phase velocities in %s:\n" : "Max/min attenuation coefficient in %s:\n",
                        ax == 0 ? "X" : "Z");
                fprintf(stdout, "   Src frequency (%lg Hz)   - slow P : %lg    %lg\n", f0, ext[b + 1][0], ext[b + 1][1]);
                fprintf(stdout, "                             fast P : %lg    %lg\n", ext[b][0], ext[b][1]);
                fprintf(stdout, "                                  S : %lg    %lg\n", ext[b + 2][0], ext[b + 2][1]);
            }
        fflush(stdout);
    }
    if (p->check_model > 1) {
        if (sm_verbose > 1) {
            fprintf(stdout, "  Saving phase velocity and attenuation ...");
            fflush(stdout);
        }
        const char *names[6] = {"qP1_x", "qP2_x", "qS_x", "qP1_z", "qP2_z", "qS_z"};
        char name[32];
        for (int k = 0; k < 6; ++k) {
            snprintf(name, sizeof(name), "vel_%s", names[k]);
            sm_write_field_nc(vel[k], name, "m/s", g->nx, g->nz, np, g->x0, g->z0, g->dx, g->dz, p->basename, 1);
        }
        for (int k = 0; k < 6; ++k) {
            snprintf(name, sizeof(name), "att_%s", names[k]);
            sm_write_field_nc(att[k], name, "1/m", g->nx, g->nz, np, g->x0, g->z0, g->dx, g->dz, p->basename, 1);
        }
    }
    for (int k = 0; k < 6; ++k) { free(vel[k]); free(att[k]); }
}

/* Checkpoint files, a layout of our own (the reference's save_checkpoint and
 * read_checkpoint1/2, io_utils.c:2355-2811, fwrite its structs and the
 * model): 8 bytes "SMPVECK", the medium (0 isotropic, 1 VTI), the
 * iteration after which it was saved, the run's sm_params, the state's size
 * and the state of sm_pve_get_state.  Written as chkpt/<base>_<it>.dat
 * (io_utils.c:2378-2383). */
static const char ck_magic[8] = "SMPVECK";

static void save_checkpoint(const sm_params *p, int vti, size_t it, sm_pve *sim)
{
    size_t bytes = 0;
    check(sm_pve_state_bytes(sim, &bytes), "sm_pve_state_bytes");
    void *blob = malloc(bytes);
    check(sm_pve_get_state(sim, blob, bytes), "sm_pve_get_state");
    struct stat sb;
    if (stat("chkpt", &sb) < 0) mkdir("chkpt", S_IRWXU);
    char fname[SM_NAME_LEN + 48];
    snprintf(fname, sizeof(fname), "chkpt/%s_%zu.dat", p->basename, it);
    FILE *fid = fopen(fname, "wb");
    if (!fid) {
        fprintf(stderr, "Error, cannot open %s for writing\n", fname);
        exit(1);
    }
    const int medium = vti;
    int ok = fwrite(ck_magic, 1, 8, fid) == 8 && fwrite(&medium, sizeof(int), 1, fid) == 1 &&
             fwrite(&it, sizeof(size_t), 1, fid) == 1 && fwrite(p, sizeof(*p), 1, fid) == 1 &&
             fwrite(&bytes, sizeof(size_t), 1, fid) == 1 && fwrite(blob, 1, bytes, fid) == bytes;
    ok = (fclose(fid) == 0) && ok;
    free(blob);
    if (!ok) {
        fprintf(stderr, "Error, cannot write %s\n", fname);
        exit(1);
    }
}

/* the parameters, the iteration and the state of a checkpoint file */
static void *read_checkpoint(const char *fname, int vti, sm_params *p, size_t *it, size_t *bytes)
{
    FILE *fid = fopen(fname, "rb");
    if (!fid) {
        fprintf(stderr, "Error, cannot open checkpoint file %s for reading\n", fname);
        exit(1);
    }
    char magic[8];
    int medium = -1;
    void *blob = NULL;
    int ok = fread(magic, 1, 8, fid) == 8 && !memcmp(magic, ck_magic, 8) &&
             fread(&medium, sizeof(int), 1, fid) == 1 && fread(it, sizeof(size_t), 1, fid) == 1 &&
             fread(p, sizeof(*p), 1, fid) == 1 && fread(bytes, sizeof(size_t), 1, fid) == 1 &&
             *bytes < ((size_t)1 << 40) && (blob = malloc(*bytes ? *bytes : 1)) != NULL &&
             fread(blob, 1, *bytes, fid) == *bytes;
    fclose(fid);
    if (!ok || medium != vti) {
        fprintf(stderr, "Error, %s is not a checkpoint of %s\n", fname, vti ? "pve_vti_pml" : "pve_iso_pml");
        exit(1);
    }
    return blob;
}

int sm_pve_program(int argc, char *argv[], int vti)
{
    sm_params p;
    sm_set_defaults(&p, 0);
    char ckfile[SM_NAME_LEN] = "";
    const int restart = sm_process_args_ck(argc, argv, &p, vti ? usage_vti : usage_iso, ckfile, sizeof(ckfile));
    size_t itstart = 0, ck_bytes = 0;
    void *ck_state = NULL;
    if (restart) {
        /* pve_iso_pml.c:525-531: the run continues after the saved iteration */
        ck_state = read_checkpoint(ckfile, vti, &p, &itstart, &ck_bytes);
        itstart++;
        if (sm_verbose) fprintf(stdout, "\n  Starting with checkpoint file %s\n", ckfile);
    }
    const size_t np = p.np;   /* 0 when abs is 0 */
    const double dt = p.dt;
    if (sm_verbose) {
        fprintf(stdout, vti ? "\n *** pve_vti - PoroViscoElastic wave propagation in 2D VTI anisotropic media ***\n\n"
                            : "\n *** pve_iso - PoroViscoElastic wave propagation in 2D isotropic media ***\n\n");
        if (sm_verbose > 1) {
            fprintf(stdout, "  Model file: \t%s\n", p.modelfile);
            fprintf(stdout, "  Source file:\t%s\n", p.sourcefile);
            fprintf(stdout, "  Output file:\t%s\n", p.outputfile);
            fprintf(stdout, "  Time window:\t%lg (s)\n", p.duration);
            fprintf(stdout, "  Time step:  \t%lg (ms)\n", p.dt * 1.e3);
        }
        fprintf(stdout, "Reading size of model ... ");
    }
    sm_grid g;
    int with_pml = 0;
    if (vti) sm_read_grid_pve_vti(p.modelfile, &g, &with_pml);
    else sm_read_grid_pve(p.modelfile, np, &g, &with_pml);
    const size_t nx2 = g.nx + 2 * np, nz2 = g.nz + 2 * np, n = nx2 * nz2;
    const int nv = vti ? V_N : 12;
    double *v[V_N];
    for (int q = 0; q < nv; ++q) v[q] = calloc(n, sizeof(double));
    if (sm_verbose) { fprintf(stdout, "done.\nReading properties of materials ... "); fflush(stdout); }
    if (vti) sm_read_model_pve_vti(p.modelfile, &g, np, with_pml, v);
    else sm_read_model_pve(p.modelfile, &g, np, with_pml, v);

    if (sm_verbose) { fprintf(stdout, "done.\nReading source parameters ... "); fflush(stdout); }
    sm_sources src;
    sm_read_source2d_np(p.sourcefile, &g, np, &src);
    const size_t ns = src.nsrc * src.ntemplate;
    const double f0 = ns ? src.s[0].f : 0.0;
    /* compute_src_fct(&src, dt, 0.5) then normalize_src_fct (src.c:43-75):
     * the functions at half steps, divided by the cell area */
    double *A = malloc((ns ? ns : 1) * sizeof(double)), *f = malloc((ns ? ns : 1) * sizeof(double));
    size_t *len = malloc((ns ? ns : 1) * sizeof(size_t)), *sidx = malloc((ns ? ns : 1) * sizeof(size_t));
    int *stype = malloc((ns ? ns : 1) * sizeof(int));
    size_t total = 0;
    for (size_t k = 0; k < ns; ++k) {
        A[k] = src.s[k].A; f[k] = src.s[k].f;
        len[k] = sm_ricker_length(f[k], 0.5 * dt);
        sidx[k] = src.s[k].idx;
        stype[k] = src.s[k].type;
        total += len[k];
    }
    double *fct = malloc((total ? total : 1) * sizeof(double));
    check(sm_compute_src_fct(ns, A, f, 0.5 * dt, len, fct), "sm_compute_src_fct");
    const double i_area = 1. / (g.dx * g.dz);
    for (size_t k = 0; k < total; ++k) fct[k] *= i_area;

    if (sm_verbose) { fprintf(stdout, "done.\nReading output parameters ... "); fflush(stdout); }
    sm_records out;
    memset(&out, 0, sizeof(out));
    snprintf(out.basename, SM_NAME_LEN, "%s", p.basename);
    sm_read_output2d_np(p.outputfile, &g, np, &out);
    sm_check_dt_trc(&out, dt);
    if (sm_verbose) fprintf(stdout, "done.\n");

    /* kinetic energy of a region (pve_iso_pml.c:758-783, 1106-1119): the
     * region must lie in the model (the reference checks it only when
     * verbose, and reads outside its arrays otherwise) */
    size_t i1E = 0, i2E = 0, j1E = 0, j2E = 0;
    FILE *efid = NULL;
    double *ecoef[6] = {NULL, NULL, NULL, NULL, NULL, NULL};
    if (p.save_energy == 1) {
        if (p.roi[0] < g.x0 || p.roi[1] > g.x0 + (double)(g.nx - 1) * g.dx || p.roi[2] < g.z0 ||
            p.roi[3] > g.z0 + (double)(g.nz - 1) * g.dz || p.roi[1] < p.roi[0] || p.roi[3] < p.roi[2]) {
            fprintf(stderr, "Region of interest for saving energy not fitting in the modeling grid\n");
            return 1;
        }
        i1E = (size_t)lround((p.roi[0] - g.x0) / g.dx);
        i2E = (size_t)lround((p.roi[1] - g.x0) / g.dx);
        j1E = (size_t)lround((p.roi[2] - g.z0) / g.dz);
        j2E = (size_t)lround((p.roi[3] - g.z0) / g.dz);
        struct stat sb;
        if (stat("E", &sb) < 0) mkdir("E", S_IRWXU);
        char fname[SM_NAME_LEN + 16];
        snprintf(fname, sizeof(fname), "E/%s_E.dat", p.basename);
        efid = fopen(fname, "w+");
        if (!efid) {
            fprintf(stderr, "Error, cannot open %s for writing\n", fname);
            return 1;
        }
        if (sm_verbose) {
            fprintf(stdout, "Saving kinetic energy as fct of time in %s\n", fname);
            fprintf(stdout, "  Region of interest is %lg/%lg/%lg/%lg\n", p.roi[0], p.roi[1], p.roi[2], p.roi[3]);
        }
        const size_t npts = (i2E - i1E + 1) * (j2E - j1E + 1);
        for (int q = 0; q < 6; ++q) ecoef[q] = malloc(npts * sizeof(double));
        for (size_t i = i1E, k = 0; i <= i2E; ++i)
            for (size_t j = j1E; j <= j2E; ++j, ++k) {
                const size_t ij = (i + np) * nz2 + j + np;
                /* densities through the reference's 1e-6 scaling and back */
                const double rs = v[RHOS][ij] * 1.e-6 * 1.e6, rf = v[RHOF][ij] * 1.e-6 * 1.e6, phi = v[PHI][ij];
                if (vti) {
                    /* eq. 7.171 of Carcione (2007): 0.5 [f1 |v|^2 - rx (vx - qx)^2
                     * - rz (vz - qz)^2 + f2 |q|^2] (pve_vti_pml.c:1112-1120) */
                    const double f1 = (1. - phi) * rs, f2 = phi * rf;
                    const double rx = (v[V_T1][ij] - 1.) / (1. / phi - 1.), rz = (v[V_T3][ij] - 1.) / (1. / phi - 1.);
                    ecoef[0][k] = f1 - rx; ecoef[1][k] = rx; ecoef[2][k] = f2 - rx;
                    ecoef[3][k] = f1 - rz; ecoef[4][k] = rz; ecoef[5][k] = f2 - rz;
                } else {
                    /* eq. 7.153-7.160 of Carcione (2007) */
                    const double r12 = -phi * rf * (v[TORT][ij] - 1.);
                    ecoef[0][k] = ecoef[3][k] = (1. - phi) * rs - r12;
                    ecoef[1][k] = ecoef[4][k] = r12;
                    ecoef[2][k] = ecoef[5][k] = phi * rf - r12;
                }
            }
    }

    if (sm_verbose) { fprintf(stdout, "Computing CPML parameters ... done.\nPreprocessing data ... "); fflush(stdout); }
    if (p.check_model > 0) {
        if (vti) check_model_vti(&p, &g, np, v, f0);
        else check_model(&p, &g, np, v, f0);
    }
    const char *dev = getenv("SEISMOD_DEVICE");
    sm_ctx *ctx = NULL;
    check(sm_ctx_create(dev ? atoi(dev) : 0, &ctx), "sm_ctx_create");
    const char *stor = getenv("SEISMOD_STORAGE");
    sm_pve_params vp;
    memset(&vp, 0, sizeof(vp));
    vp.nx = g.nx; vp.nz = g.nz; vp.np = np; vp.dx = g.dx; vp.dz = g.dz; vp.dt = dt;
    vp.vmax = p.vmax; vp.pml_order = p.pml_order; vp.Rc = p.Rc; vp.kappa_max = p.kappa_max;
    vp.alpha_max = p.alpha_max; vp.f0 = f0; vp.abs = p.iwipe;
    vp.storage = (stor && !strcmp(stor, "f32")) ? SM_STORE_F32 : SM_STORE_F64;
    /* pml_median 1 smooths the isotropic medium's layers (pve_iso_pml.c:833-845);
     * pve_vti_pml reads the keyword and does nothing with it */
    if (p.pml_median == 1 && !vti) vp.flags |= SM_PVE_PML_MEDIAN;
    sm_pve *sim = NULL;
    if (vti)
        check(sm_pve_create_vti(ctx, &vp, v[V_KM], v[V_KS], v[V_KF], v[V_PHI], v[V_MU], v[V_RHOS], v[V_RHOF],
                                v[V_T1], v[V_T3], v[V_ETA], v[V_K1], v[V_K3], v[V_Q], v[V_FQ], v[V_EPS],
                                v[V_DELTA], &sim),
              "sm_pve_create_vti");
    else
        check(sm_pve_create(ctx, &vp, v[KM], v[KS], v[KF], v[PHI], v[MU], v[RHOS], v[RHOF], v[TORT], v[ETA],
                            v[KAPPA], v[QQ], v[FQ], &sim),
              "sm_pve_create");
    check(sm_pve_set_sources(sim, ns, stype, sidx, len, fct), "sm_pve_set_sources");
    if (sm_verbose) fprintf(stdout, "done.\n");

    const size_t nsteps = sm_nsteps(p.duration, dt);   /* pve_iso_pml.c:1243-1244 */
    if (efid)
        check(sm_pve_set_energy_axes(sim, i1E + np, i2E + np, j1E + np, j2E + np, ecoef[0], ecoef[1], ecoef[2],
                                     ecoef[3], ecoef[4], ecoef[5], nsteps),
              "sm_pve_set_energy_axes");
    /* records of other components write nothing (pve_iso_pml.c:1330-1389) */
    size_t ntr = 0;
    for (size_t r = 0; r < out.nrec; ++r) ntr += out.r[r].type == REC_TRACE && recorded(out.r[r].comp, vti);
    size_t *ridx = malloc((ntr ? ntr : 1) * sizeof(size_t));
    int *rcomp = malloc((ntr ? ntr : 1) * sizeof(int));
    size_t *rmap = malloc((ntr ? ntr : 1) * sizeof(size_t));
    for (size_t r = 0, q = 0; r < out.nrec; ++r)
        if (out.r[r].type == REC_TRACE && recorded(out.r[r].comp, vti)) {
            ridx[q] = out.r[r].idx; rcomp[q] = out.r[r].comp; rmap[q] = r; ++q;
            if (p.segy == 1) {   /* pve_iso_pml.c:719-726 */
                out.r[r].nsamples = (size_t)round(p.duration / out.r[r].dt);
                out.r[r].data = calloc(out.r[r].nsamples + 1, sizeof(float));
            }
        }
    check(sm_pve_set_receivers(sim, ntr, ridx, rcomp, nsteps), "sm_pve_set_receivers");

    if (ck_state) {
        check(sm_pve_set_state(sim, ck_state, ck_bytes), "sm_pve_set_state");
        free(ck_state);
    }

    if (sm_verbose) {
        fprintf(stdout, "Starting main loop (%zu iterations)\n", nsteps);
        fflush(stdout);
    }
    double *field = malloc(n * sizeof(double));
    size_t done = itstart;
    for (size_t it = itstart; it < nsteps; ++it) {
        const double t = dt * (it + 1);
        if ((it + 1) % 50 == 0 && sm_verbose) {
            fprintf(stdout, "  Iteration %zu, t = %g s\n", it + 1, t);
            fflush(stdout);
        }
        int snap = 0;
        for (size_t r = 0; r < out.nrec; ++r)
            if (out.r[r].type == REC_SNAPSHOT && recorded(out.r[r].comp, vti) &&
                sm_record_due(it, dt, out.r[r].t0, out.r[r].dt))
                snap = 1;
        /* pve_iso_pml.c:1406-1411 */
        const int ck = p.chkpt_inc > 0 && it > 0 && it % (size_t)p.chkpt_inc == 0;
        if (!snap && !ck && it + 1 < nsteps) continue;
        check(sm_pve_run(sim, it + 1 - done), "sm_pve_run");
        done = it + 1;
        if (ck) {
            if (sm_verbose > 1) printf("Saving checkpoint data at iteration %zu\n", it);
            save_checkpoint(&p, vti, it, sim);
        }
        for (size_t r = 0; r < out.nrec; ++r)
            if (out.r[r].type == REC_SNAPSHOT && recorded(out.r[r].comp, vti) &&
                sm_record_due(it, dt, out.r[r].t0, out.r[r].dt)) {
                if (sm_verbose > 1) printf("Writing snapshot, t = %g s\n", t);
                check(sm_pve_get_field(sim, out.r[r].comp, field), "sm_pve_get_field");
                sm_write_snapshot_nc(field, t, g.nx, g.nz, np, g.x0, g.z0, g.dx, g.dz, out.basename, r,
                                     p.plot_strips);
            }
    }
    if (ntr) {
        double *tr = malloc(nsteps * ntr * sizeof(double));
        check(sm_pve_get_traces(sim, 0, nsteps, tr), "sm_pve_get_traces");
        for (size_t it = 0; it < nsteps; ++it)
            for (size_t q = 0; q < ntr; ++q) {
                sm_record *r = &out.r[rmap[q]];
                const double t = dt * (it + 1);
                if (!sm_record_due(it, dt, r->t0, r->dt)) continue;
                if (p.segy == 1) {
                    /* fill_data_segy (io_utils.c:1770-1774) */
                    const size_t s = (size_t)round(t / r->dt);
                    if (s < r->nsamples) r->data[s] = (float)tr[it * ntr + q];
                } else {
                    sm_write_trace(&out, rmap[q], t, tr[it * ntr + q]);
                }
            }
        free(tr);
    }
    if (efid) {
        double *E = malloc((nsteps ? nsteps : 1) * sizeof(double));
        check(sm_pve_get_energy(sim, 0, nsteps, E), "sm_pve_get_energy");
        for (size_t it = 0; it < nsteps; ++it) fprintf(efid, "%le   %le\n", dt * (it + 1), E[it]);
        fclose(efid);
        free(E);
    }
    if (p.segy == 1) sm_save_segy(&p, &out, &src);
    sm_close_traces(&out);
    check(sm_pve_destroy(sim), "sm_pve_destroy");
    check(sm_ctx_destroy(ctx), "sm_ctx_destroy");
    for (int q = 0; q < nv; ++q) free(v[q]);
    free(A); free(f); free(len); free(sidx); free(stype); free(fct); free(ridx); free(rcomp); free(rmap);
    free(field);
    for (int q = 0; q < 6; ++q) free(ecoef[q]);
    sm_free_sources(&src);
    sm_free_records(&out);
    if (sm_verbose) fprintf(stdout, "\nEnd of computation\n");
    return 0;
}
