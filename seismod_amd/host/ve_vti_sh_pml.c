/*
 * ve_vti_sh_pml -- viscoelastic SH-wave propagation in 2D VTI media
 * (pseudo-spectral derivatives, CPML, fourth-order Runge-Kutta), running
 * the time loop on an MI355X through libseismod_hip.
 *
 * Drop-in for seismod's ve_vti_sh_pml: same command line, .par keywords,
 * model ("Vs0 rho gamma Q f" per node), source (Sxy, Syz; 1- or 9-point
 * template) and record (Vy, txy, tyz) files, and the same trc/, ssh/ and
 * segy/ outputs (ve_vti_sh_pml.c:236-807).  Set-up and update arithmetic
 * live in the library (sm_sh_*); this program reads the inputs, drives the
 * time loop and writes the records.
 *
 * Environment: SEISMOD_STORAGE=f32 runs the FP32 kernels; the default is
 * FP64 (the reference's precision).  SEISMOD_DEVICE selects the GPU.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "seismod_hip.h"
#include "sm_io.h"

static const char *usage_text =
    "Format of parameter file:\n\n"
    "One line per parameter, first word is the value of the parameter,\n"
    "next is a keyword comprise between a # and a comma, and a comment\n"
    "can be added after the comma, i.e we have on each line\n\n"
    "value     # keyword, optional comment\n\n"
    "Keywords: model, source, output, basename, time (s), dt (ms), segy, shotpoint,\n"
    "simulateCMP, abs, nl, plstr, kappa_max, alpha_max, Rc\n\n"
    "Model file: nx nz / dx dz / x0 z0 / vti_sh_viscoelastic / then one\n"
    "\"Vs0 rho gamma Q f\" line per node, x outer and z inner (one column only:\n"
    "laterally uniform)\n\n"
    "Source file: nsrc / template points (1 or 9) / then per source: type\n"
    "(Sxy, Syz), strength, frequency [Hz], x z [m]\n\n"
    "Record file: nrec / then per record: component (Vy, txy, tyz),\n"
    "Trace x z t0[s] dt[ms] or Snapshot t0[s] dt[ms]\n";

static void check(int st, const char *where)
{
    if (st != SM_OK) {
        fprintf(stderr, "Error in %s: %s (%s)\n", where, sm_strerror(st), sm_last_error());
        exit(1);
    }
}

static int recorded(int c) { return c == SM_VY || c == SM_TXY || c == SM_TYZ; }

int main(int argc, char *argv[])
{
    sm_params p;
    sm_set_defaults(&p, 0);
    sm_process_args(argc, argv, &p, usage_text);
    const size_t np = p.np;   /* 0 when abs is 0 */
    const double dt = p.dt;
    if (sm_verbose) {
        fprintf(stdout, "\n *** ve_vti_sh - ViscoElastic SH-wave propagation in 2D VTI anisotropic media ***\n\n");
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
    sm_read_grid_ve_sh(p.modelfile, &g);
    const size_t nx2 = g.nx + 2 * np, nz2 = g.nz + 2 * np, n = nx2 * nz2;
    double *mu = calloc(n, sizeof(double)), *rho = calloc(n, sizeof(double)),
           *gamma = calloc(n, sizeof(double)), *Q = calloc(n, sizeof(double)),
           *fr = calloc(n, sizeof(double));
    if (sm_verbose) { fprintf(stdout, "done.\nReading properties of materials ... "); fflush(stdout); }
    sm_read_model_ve_sh(p.modelfile, &g, np, mu, rho, gamma, Q, fr);

    if (sm_verbose) { fprintf(stdout, "done.\nReading source parameters ... "); fflush(stdout); }
    sm_sources src;
    sm_read_source2d_np(p.sourcefile, &g, np, &src);
    const size_t ns = src.nsrc * src.ntemplate;
    /* compute_src_fct(&src, dt, 0.5) then normalize_src_fct (src.c:43-75):
     * the functions at half steps, divided by the cell area */
    double *A = malloc(ns * sizeof(double)), *f = malloc(ns * sizeof(double));
    size_t *len = malloc(ns * sizeof(size_t)), *sidx = malloc(ns * sizeof(size_t));
    int *stype = malloc(ns * sizeof(int));
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
    if (sm_verbose) fprintf(stdout, "done.\nComputing CPML parameters ... ");

    const char *dev = getenv("SEISMOD_DEVICE");
    sm_ctx *ctx = NULL;
    check(sm_ctx_create(dev ? atoi(dev) : 0, &ctx), "sm_ctx_create");
    const char *stor = getenv("SEISMOD_STORAGE");
    sm_sh_params sp;
    memset(&sp, 0, sizeof(sp));
    sp.nx = g.nx; sp.nz = g.nz; sp.np = np; sp.dx = g.dx; sp.dz = g.dz; sp.dt = dt;
    sp.vmax = p.vmax; sp.pml_order = p.pml_order; sp.Rc = p.Rc; sp.kappa_max = p.kappa_max;
    sp.alpha_max = p.alpha_max; sp.f0 = ns ? src.s[0].f : 0.0; sp.abs = p.iwipe;
    sp.storage = (stor && !strcmp(stor, "f32")) ? SM_STORE_F32 : SM_STORE_F64;
    sm_sh *sim = NULL;
    check(sm_sh_create(ctx, &sp, mu, rho, gamma, Q, fr, &sim), "sm_sh_create");
    check(sm_sh_set_sources(sim, ns, stype, sidx, len, fct), "sm_sh_set_sources");
    if (sm_verbose) fprintf(stdout, "done.\n");

    const size_t nsteps = sm_nsteps(p.duration, dt);   /* ve_vti_sh_pml.c:630-631 */
    /* records of other components write nothing (ve_vti_sh_pml.c:698-731) */
    size_t ntr = 0;
    for (size_t r = 0; r < out.nrec; ++r) ntr += out.r[r].type == REC_TRACE && recorded(out.r[r].comp);
    size_t *ridx = malloc((ntr ? ntr : 1) * sizeof(size_t));
    int *rcomp = malloc((ntr ? ntr : 1) * sizeof(int));
    size_t *rmap = malloc((ntr ? ntr : 1) * sizeof(size_t));
    for (size_t r = 0, q = 0; r < out.nrec; ++r)
        if (out.r[r].type == REC_TRACE && recorded(out.r[r].comp)) {
            ridx[q] = out.r[r].idx; rcomp[q] = out.r[r].comp; rmap[q] = r; ++q;
            if (p.segy == 1) {   /* ve_vti_sh_pml.c:413-420 */
                out.r[r].nsamples = (size_t)round(p.duration / out.r[r].dt);
                out.r[r].data = calloc(out.r[r].nsamples + 1, sizeof(float));
            }
        }
    check(sm_sh_set_receivers(sim, ntr, ridx, rcomp, nsteps), "sm_sh_set_receivers");

    if (sm_verbose) {
        fprintf(stdout, "Starting main loop (%zu iterations)\n", nsteps);
        fflush(stdout);
    }
    double *field = malloc(n * sizeof(double));
    size_t done = 0;
    for (size_t it = 0; it < nsteps; ++it) {
        const double t = dt * (it + 1);
        if ((it + 1) % 50 == 0 && sm_verbose) {
            fprintf(stdout, "  Iteration %zu, t = %g s\n", it + 1, t);
            fflush(stdout);
        }
        int snap = 0;
        for (size_t r = 0; r < out.nrec; ++r)
            if (out.r[r].type == REC_SNAPSHOT && recorded(out.r[r].comp) &&
                sm_record_due(it, dt, out.r[r].t0, out.r[r].dt))
                snap = 1;
        if (!snap && it + 1 < nsteps) continue;
        check(sm_sh_run(sim, it + 1 - done), "sm_sh_run");
        done = it + 1;
        for (size_t r = 0; r < out.nrec; ++r)
            if (out.r[r].type == REC_SNAPSHOT && recorded(out.r[r].comp) &&
                sm_record_due(it, dt, out.r[r].t0, out.r[r].dt)) {
                if (sm_verbose > 1) printf("Writing snapshot, t = %g s\n", t);
                check(sm_sh_get_field(sim, out.r[r].comp, field), "sm_sh_get_field");
                sm_write_snapshot_nc(field, t, g.nx, g.nz, np, g.x0, g.z0, g.dx, g.dz, out.basename, r,
                                     p.plot_strips);
            }
    }
    if (ntr) {
        double *tr = malloc(nsteps * ntr * sizeof(double));
        check(sm_sh_get_traces(sim, 0, nsteps, tr), "sm_sh_get_traces");
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
    if (p.segy == 1) sm_save_segy(&p, &out, &src);
    sm_close_traces(&out);
    check(sm_sh_destroy(sim), "sm_sh_destroy");
    check(sm_ctx_destroy(ctx), "sm_ctx_destroy");
    free(mu); free(rho); free(gamma); free(Q); free(fr); free(A); free(f); free(len); free(sidx);
    free(stype); free(fct); free(ridx); free(rcomp); free(rmap); free(field);
    sm_free_sources(&src);
    sm_free_records(&out);
    if (sm_verbose) fprintf(stdout, "\nEnd of computation\n");
    return 0;
}
