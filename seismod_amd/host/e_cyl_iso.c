/*
 * e_cyl_iso -- elastic wave propagation in 2.5D isotropic media, on a
 * staggered grid in cylindrical coordinates with a vertical symmetry axis,
 * running the time loop on an MI355X through libseismod_hip.
 *
 * Drop-in for seismod's e_cyl_iso: same command line, .par keywords
 * (including azimuthal_mode), model (isotropic_elastic_cyl, r fastest),
 * source and record files, and the same trc/ and ssh/ outputs
 * (e_cyl_iso.c:246-885).  Set-up and update arithmetic live in the library
 * (sm_cyl_*); this program reads the inputs, drives the time loop and
 * writes the records.
 *
 * Environment: SEISMOD_STORAGE=f32 runs FP32 storage and arithmetic; the
 * default is FP64 (the reference's operation order, bit for bit).
 * SEISMOD_DEVICE selects the GPU.
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
    "next is a keyword comprised between a # and a comma, and a comment\n"
    "can be added after the comma, i.e we have on each line\n\n"
    "value     # keyword, optional comment\n\n"
    "Keywords: model, source, output, basename, time (s), dt (ms), abs, nl, plstr,\n"
    "azimuthal_mode, Rc, kappa_max, alpha_max\n\n"
    "Model file: nr nz / dr dz / r0 z0 / isotropic_elastic_cyl / then one\n"
    "\"Vp Vs rho\" line per node, r fastest (one r profile only: copied along z)\n\n"
    "Source file: nsrc / template points (1 or 9) / then per source: type\n"
    "(Kurkjian, Fr, Fz, Ftheta), strength (peak volume change in m^3 for\n"
    "Kurkjian), frequency [Hz], r z [m]\n\n"
    "Record file: nrec / then per record: component (Vr, Vtheta, Vz, trr, trtheta,\n"
    "trz, tthetatheta, tthetaz, tzz, div, curl), Trace r z t0[s] dt[ms] or\n"
    "Snapshot t0[s] dt[ms]\n";

static void check(int st, const char *where)
{
    if (st != SM_OK) {
        fprintf(stderr, "Error in %s: %s (%s)\n", where, sm_strerror(st), sm_last_error());
        if (st == SM_ERR_CFL) {
            fprintf(stderr, "\n *** Warning!  Time step too large, aborting ***\n");
            abort();   /* e_cyl_iso.c:454-458 */
        }
        if (st == SM_ERR_ARG && strstr(sm_last_error(), "not finite")) abort();   /* 417-452 */
        exit(1);
    }
}

/* divergence and curl of the velocity (propagate.c:4090-4119) for
 * snapshots, on the host from the three velocity fields */
static void div_curl(const double *vr, const double *vt, const double *vz, size_t nx2, size_t nz2,
                     double dx, double dz, int n, double *d, double *c)
{
    if (d)
        for (size_t i = 0; i + 1 < nx2; ++i)
            for (size_t j = 0; j + 1 < nz2; ++j)
                d[i * nz2 + j] = ((i + 1) * vr[(i + 1) * nz2 + j] - i * vr[i * nz2 + j]) / (i + 0.5) +
                                 n * vt[i * nz2 + j] / ((i + 0.5) * dx) +
                                 (vz[i * nz2 + j + 1] - vz[i * nz2 + j]) / dz;
    if (c)
        for (size_t i = (n % 2) ? 0 : 1; i < nx2; ++i)
            for (size_t j = 1; j < nz2; ++j)
                c[i * nz2 + j] = (vr[i * nz2 + j] - vr[i * nz2 + j - 1]) / dz -
                                 (i == 0 ? (vz[j] - -vz[j]) : (vz[i * nz2 + j] - vz[(i - 1) * nz2 + j])) / dx;
}

int main(int argc, char *argv[])
{
    sm_params p;
    sm_set_defaults(&p, 0);
    sm_process_args(argc, argv, &p, usage_text);
    const size_t np = p.np;
    const double dt = p.dt;
    if (p.mode < 0) {
        fprintf(stderr, "Error: azimuthal mode must be >= 0\n");
        return 1;
    }
    if (sm_verbose) {
        fprintf(stdout, "\n *** e_cyl_iso - Elastic wave propagation in 2.5D isotropic media in "
                        "cylindrical coordinates ***\n\n");
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
    const size_t nx2 = g.nx + np, nz2 = g.nz + 2 * np, nn = nx2 * nz2;
    double *vp = calloc(nn, sizeof(double)), *vs = calloc(nn, sizeof(double)),
           *rho = calloc(nn, sizeof(double));
    if (sm_verbose) { fprintf(stdout, "done.\nReading properties of materials ... "); fflush(stdout); }
    sm_read_model_e_cyl(p.modelfile, &g, np, vp, vs, rho);

    if (sm_verbose) { fprintf(stdout, "done.\nReading source parameters ... "); fflush(stdout); }
    sm_sources src;
    sm_read_source_cyl(p.sourcefile, &g, np, &src);
    const size_t ns = src.nsrc * src.ntemplate;
    if (ns == 0) {
        fprintf(stderr, "Error: no source in %s\n", p.sourcefile);
        return 1;
    }
    const int kurk = src.s[0].type == SM_KURKJIAN;
    for (size_t k = 1; k < ns; ++k)
        if (src.s[k].type == SM_KURKJIAN || kurk) {
            /* compute_kurkjian fills source 0 only and add_kurkjian_src
             * injects source 0 only (src.c:202-238): other sources beside a
             * Kurkjian one have no wavelet in the reference */
            fprintf(stderr, "Error: a Kurkjian source must be the only source\n");
            return 1;
        }

    const char *dev = getenv("SEISMOD_DEVICE");
    sm_ctx *ctx = NULL;
    check(sm_ctx_create(dev ? atoi(dev) : 0, &ctx), "sm_ctx_create");
    const char *stor = getenv("SEISMOD_STORAGE");
    sm_cyl_params cp;
    memset(&cp, 0, sizeof(cp));
    cp.nx = g.nx; cp.nz = g.nz; cp.np = np; cp.dr = g.dx; cp.dz = g.dz; cp.dt = dt;
    cp.mode = p.mode;
    cp.vmax = p.vmax; cp.pml_order = p.pml_order; cp.Rc = p.Rc; cp.kappa_max = p.kappa_max;
    cp.alpha_max = p.alpha_max; cp.f0 = src.s[0].f;
    cp.storage = (stor && !strcmp(stor, "f32")) ? SM_STORE_F32 : SM_STORE_F64;
    sm_cyl *sim = NULL;
    check(sm_cyl_create(ctx, &cp, vp, vs, rho, &sim), "sm_cyl_create");

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
    if (kurk) {
        /* e_cyl_iso.c:470-474, compute_kurkjian with lambda at the source */
        const size_t c = sidx[0];
        const double lam = rho[c] * (vp[c] * vp[c] - 2. * vs[c] * vs[c]);
        check(sm_kurkjian_fct(lam, g.dx, g.dz, src.s[0].x, p.mode, A[0], f[0], dt, len[0], fct),
              "sm_kurkjian_fct");
    } else {
        check(sm_compute_src_fct(ns, A, f, dt, len, fct), "sm_compute_src_fct");
    }
    check(sm_cyl_set_sources(sim, ns, stype, sidx, len, fct), "sm_cyl_set_sources");
    if (sm_verbose) fprintf(stdout, "done.\nApplying PML absorbing boundaries (%zu layers) ", np);

    if (sm_verbose) { fprintf(stdout, "Reading output parameters ... "); fflush(stdout); }
    sm_records out;
    memset(&out, 0, sizeof(out));
    snprintf(out.basename, SM_NAME_LEN, "%s", p.basename);
    sm_read_output_cyl(p.outputfile, &g, np, &out);
    sm_check_dt_trc(&out, dt);
    if (sm_verbose) fprintf(stdout, "done.\n");

    const size_t nsteps = sm_nsteps(p.duration, dt);   /* e_cyl_iso.c:586-587 */
    size_t ntr = 0;
    for (size_t r = 0; r < out.nrec; ++r) {
        const int c = out.r[r].comp;
        if (c != SM_VR && c != SM_VT && c != SM_VZ && c != SM_TRR && c != SM_TRT && c != SM_TRZ &&
            c != SM_TTT && c != SM_TTZ && c != SM_TZZ && c != SM_DIV && c != SM_CURL) {
            fprintf(stderr, "Error: record %zu: e_cyl_iso records Vr, Vtheta, Vz, trr, trtheta, trz, "
                            "tthetatheta, tthetaz, tzz, div or curl\n", r + 1);
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
    check(sm_cyl_set_receivers(sim, ntr, ridx, rcomp, nsteps), "sm_cyl_set_receivers");

    if (sm_verbose) {
        fprintf(stdout, "Starting main loop (%zu iterations)\n", nsteps);
        fflush(stdout);
    }
    double *field = malloc(nn * sizeof(double)), *v3 = NULL;
    size_t done = 0;
    for (size_t it = 0; it < nsteps; ++it) {
        const double t = dt * (it + 1);
        if ((it + 1) % 50 == 0 && sm_verbose) {
            fprintf(stdout, "  Iteration %zu, t = %g s\n", it + 1, t);
            fflush(stdout);
        }
        int snap = 0;
        for (size_t r = 0; r < out.nrec; ++r)
            if (out.r[r].type == REC_SNAPSHOT && sm_record_due(it, dt, out.r[r].t0, out.r[r].dt))
                snap = 1;
        if (!snap && it + 1 < nsteps) continue;
        check(sm_cyl_run(sim, it + 1 - done), "sm_cyl_run");
        done = it + 1;
        for (size_t r = 0; r < out.nrec; ++r) {
            if (out.r[r].type != REC_SNAPSHOT || !sm_record_due(it, dt, out.r[r].t0, out.r[r].dt))
                continue;
            if (sm_verbose > 1) printf("Writing snapshot, t = %g s\n", t);
            const int c = out.r[r].comp;
            if (c == SM_DIV || c == SM_CURL) {
                if (!v3) v3 = malloc(3 * nn * sizeof(double));
                check(sm_cyl_get_field(sim, SM_VR, v3), "sm_cyl_get_field");
                check(sm_cyl_get_field(sim, SM_VT, v3 + nn), "sm_cyl_get_field");
                check(sm_cyl_get_field(sim, SM_VZ, v3 + 2 * nn), "sm_cyl_get_field");
                memset(field, 0, nn * sizeof(double));
                div_curl(v3, v3 + nn, v3 + 2 * nn, nx2, nz2, g.dx, g.dz, p.mode,
                         c == SM_DIV ? field : NULL, c == SM_CURL ? field : NULL);
            } else {
                check(sm_cyl_get_field(sim, c, field), "sm_cyl_get_field");
            }
            sm_write_snapshot_cyl_nc(field, t, g.nx, g.nz, np, g.x0, g.z0, g.dx, g.dz, out.basename,
                                     r, p.plot_strips);
        }
    }
    if (ntr) {
        double *tr = malloc(nsteps * ntr * sizeof(double));
        check(sm_cyl_get_traces(sim, 0, nsteps, tr), "sm_cyl_get_traces");
        for (size_t it = 0; it < nsteps; ++it)
            for (size_t q = 0; q < ntr; ++q) {
                const sm_record *r = &out.r[rmap[q]];
                if (sm_record_due(it, dt, r->t0, r->dt))
                    sm_write_trace(&out, rmap[q], dt * (it + 1), tr[it * ntr + q]);
            }
        free(tr);
    }
    sm_close_traces(&out);
    check(sm_cyl_destroy(sim), "sm_cyl_destroy");
    check(sm_ctx_destroy(ctx), "sm_ctx_destroy");
    free(vp); free(vs); free(rho); free(A); free(f); free(len); free(sidx); free(stype); free(fct);
    free(ridx); free(rcomp); free(rmap); free(field); free(v3);
    sm_free_sources(&src);
    sm_free_records(&out);
    if (sm_verbose) fprintf(stdout, "\nEnd of computation\n");
    return 0;
}
