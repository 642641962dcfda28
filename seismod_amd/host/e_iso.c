/*
 * e_iso -- 2D elastic wave propagation in isotropic media (staggered grid,
 * split PML), running the time loop on an MI355X through libseismod_hip.
 *
 * Drop-in for seismod's e_iso: same command line, .par keywords, model,
 * source and record files, and the same trc/ and ssh/ outputs
 * (e_iso.c:290-699).  Set-up and update arithmetic live in the library
 * (sm_e2d_*); this program reads the inputs, drives the time loop and
 * writes the records.
 *
 * Environment: SEISMOD_STORAGE=f64 runs the FP64 kernels (same operation
 * order as the reference); the default is FP32.  SEISMOD_DEVICE selects the GPU.
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
    "Keywords: model, source, output, basename, time (s), dt (ms), abs, nl, plstr\n\n"
    "Model file: nx nz / dx dz / x0 z0 / isotropic_elastic / then one \"Vp Vs rho\"\n"
    "line per node, x outer and z inner (one column only: laterally uniform)\n\n"
    "Source file: nsrc / template points (1 or 9) / then per source: type\n"
    "(Sx, Sz, Sxz, Bulk), strength, frequency [Hz], x z [m]\n\n"
    "Record file: nrec / then per record: component (Vx, Vz, txx, tzz, txz),\n"
    "Trace x z t0[s] dt[ms] or Snapshot t0[s] dt[ms]\n";

static void check(int st, const char *where)
{
    if (st != SM_OK) {
        fprintf(stderr, "Error in %s: %s (%s)\n", where, sm_strerror(st), sm_last_error());
        if (st == SM_ERR_CFL) {
            fprintf(stderr, "\n *** Warning!  Time step too large, aborting ***\n");
            abort();   /* e_iso.c:476-479 */
        }
        exit(1);
    }
}

int main(int argc, char *argv[])
{
    sm_params p;
    sm_set_defaults(&p, 0);
    sm_process_args(argc, argv, &p, usage_text);
    const size_t npml = p.np;   /* e_iso.c:346-347: PML kept outside the grid */
    const double dt = p.dt;
    if (sm_verbose) {
        fprintf(stdout, "\n *** e_iso - Elastic wave propagation in 2D isotropic media ***\n\n");
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
    sm_read_grid2d(p.modelfile, &g);
    if (g.dx != g.dz) {
        fprintf(stderr, "Error: grid cell size must be equal in X and Z, aborting.\n");
        abort();   /* e_iso.c:363-366 */
    }
    if (npml == 0) {
        /* e_iso.c with abs 0 indexes its PML strips with Npml-1 = SIZE_MAX */
        fprintf(stderr, "Error: e_iso needs absorbing layers (abs 1, nl >= 1)\n");
        return 1;
    }
    const size_t n = g.nx * g.nz;
    double *vp = malloc(n * sizeof(double)), *vs = malloc(n * sizeof(double)),
           *rho = malloc(n * sizeof(double));
    if (sm_verbose) { fprintf(stdout, "done.\nReading properties of materials ... "); fflush(stdout); }
    sm_read_model_e(p.modelfile, &g, vp, vs, rho);

    const char *dev = getenv("SEISMOD_DEVICE");
    sm_ctx *ctx = NULL;
    check(sm_ctx_create(dev ? atoi(dev) : 0, &ctx), "sm_ctx_create");
    const char *stor = getenv("SEISMOD_STORAGE");
    sm_e2d_params ep;
    memset(&ep, 0, sizeof(ep));
    ep.nx = g.nx; ep.nz = g.nz; ep.npml = npml; ep.h = g.dx; ep.dt = dt;
    ep.storage = (stor && !strcmp(stor, "f64")) ? SM_STORE_F64 : SM_STORE_F32;
    ep.q1_compat = 1;
    ep.damping_scale = 1.0;
    sm_e2d *sim = NULL;
    check(sm_e2d_create(ctx, &ep, vp, vs, rho, &sim), "sm_e2d_create");
    if (sm_verbose) fprintf(stdout, "done.\nApplying PML absorbing boundaries (%zu layers)", npml);

    if (sm_verbose) { fprintf(stdout, "\nReading source parameters ... "); fflush(stdout); }
    sm_sources src;
    sm_read_source2d(p.sourcefile, &g, &src);
    const size_t ns = src.nsrc * src.ntemplate;
    double *A = malloc(ns * sizeof(double)), *f = malloc(ns * sizeof(double));
    size_t *len = malloc(ns * sizeof(size_t)), *sidx = malloc(ns * sizeof(size_t));
    int *stype = malloc(ns * sizeof(int));
    size_t total = 0;
    for (size_t k = 0; k < ns; ++k) {
        A[k] = src.s[k].A; f[k] = src.s[k].f;
        len[k] = sm_ricker_length(f[k], dt);
        sidx[k] = src.s[k].idx;
        stype[k] = src.s[k].type;
        total += len[k];
    }
    double *fct = malloc((total ? total : 1) * sizeof(double));
    check(sm_compute_src_fct(ns, A, f, dt, len, fct), "sm_compute_src_fct");
    check(sm_e2d_set_sources(sim, ns, stype, sidx, len, fct), "sm_e2d_set_sources");

    if (sm_verbose) { fprintf(stdout, "done.\nReading output parameters ... "); fflush(stdout); }
    sm_records out;
    memset(&out, 0, sizeof(out));
    snprintf(out.basename, SM_NAME_LEN, "%s", p.basename);
    sm_read_output2d(p.outputfile, &g, &out);
    sm_check_dt_trc(&out, dt);
    if (sm_verbose) fprintf(stdout, "done.\n");

    const size_t nsteps = sm_nsteps(p.duration, dt);   /* e_iso.c:575-576 */
    /* traces: every record is read back after each step; the record
     * predicate of e_iso.c:643-658 picks the samples written */
    size_t ntr = 0;
    for (size_t r = 0; r < out.nrec; ++r) {
        const int c = out.r[r].comp;
        if (c != SM_VX && c != SM_VZ && c != SM_TXX && c != SM_TZZ && c != SM_TXZ) {
            fprintf(stderr, "Error: record %zu: e_iso records Vx, Vz, txx, tzz or txz\n", r + 1);
            return 1;
        }
        ntr += out.r[r].type == REC_TRACE;
    }
    size_t *ridx = malloc((ntr ? ntr : 1) * sizeof(size_t));
    int *rcomp = malloc((ntr ? ntr : 1) * sizeof(int));
    size_t *rmap = malloc((ntr ? ntr : 1) * sizeof(size_t));
    for (size_t r = 0, q = 0; r < out.nrec; ++r)
        if (out.r[r].type == REC_TRACE) {
            ridx[q] = out.r[r].idx; rcomp[q] = out.r[r].comp; rmap[q] = r; ++q;
        }
    check(sm_e2d_set_receivers(sim, ntr, ridx, rcomp, nsteps), "sm_e2d_set_receivers");

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
        /* snapshots due after this step need the field at this step */
        int snap = 0;
        for (size_t r = 0; r < out.nrec; ++r)
            if (out.r[r].type == REC_SNAPSHOT && sm_record_due(it, dt, out.r[r].t0, out.r[r].dt))
                snap = 1;
        if (!snap && it + 1 < nsteps) continue;
        check(sm_e2d_run(sim, it + 1 - done), "sm_e2d_run");
        done = it + 1;
        for (size_t r = 0; r < out.nrec; ++r)
            if (out.r[r].type == REC_SNAPSHOT && sm_record_due(it, dt, out.r[r].t0, out.r[r].dt)) {
                if (sm_verbose > 1)
                    printf("Writing snapshot, t = %g s\n", t);
                check(sm_e2d_get_field(sim, out.r[r].comp, field), "sm_e2d_get_field");
                sm_write_snapshot_nc(field, t, g.nx, g.nz, 0, g.x0, g.z0, g.dx, g.dz, out.basename,
                                     r, p.plot_strips);
            }
    }
    if (ntr) {
        double *tr = malloc(nsteps * ntr * sizeof(double));
        check(sm_e2d_get_traces(sim, 0, nsteps, tr), "sm_e2d_get_traces");
        for (size_t it = 0; it < nsteps; ++it)
            for (size_t q = 0; q < ntr; ++q) {
                const sm_record *r = &out.r[rmap[q]];
                if (sm_record_due(it, dt, r->t0, r->dt))
                    sm_write_trace(&out, rmap[q], dt * (it + 1), tr[it * ntr + q]);
            }
        free(tr);
    }
    sm_close_traces(&out);
    check(sm_e2d_destroy(sim), "sm_e2d_destroy");
    check(sm_ctx_destroy(ctx), "sm_ctx_destroy");
    free(vp); free(vs); free(rho); free(A); free(f); free(len); free(sidx); free(stype); free(fct);
    free(ridx); free(rcomp); free(rmap); free(field);
    sm_free_sources(&src);
    sm_free_records(&out);
    if (sm_verbose) fprintf(stdout, "\nEnd of computation\n");
    return 0;
}
