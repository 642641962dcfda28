/*
 * ve_vti_pml -- viscoelastic P-SV wave propagation in 2D VTI media
 * (pseudo-spectral derivatives, CPML, fourth-order Runge-Kutta, L
 * quasi-dilatational relaxation mechanisms and one shear mechanism), running
 * the time loop on an MI355X through libseismod_hip.
 *
 * Drop-in for seismod's ve_vti_pml: same command line, .par keywords,
 * model ("Vp0 Vs0 rho epsilon delta", L pairs "Q1 f1", "Q2 f2" per node),
 * source (Sx, Sz, Bulk, Sxz; 1- or 9-point template) and record (Vx, Vz,
 * txx, tzz, txz) files, and the same outputs: the moduli files
 * <base>_c11.nc ... <base>_K0.nc, trc/, ssh/ and segy/ (ve_vti_pml.c:250-913).
 * Set-up and update arithmetic live in the library (sm_ve_*); this program
 * reads the inputs, drives the time loop and writes the records.
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
    "Model file: nx nz / dx dz / x0 z0 / vti_viscoelastic / L / then one line per\n"
    "node, x outer and z inner: Vp0 Vs0 rho epsilon delta, L pairs Q1 f1, Q2 f2\n"
    "(one column only: laterally uniform)\n\n"
    "Source file: nsrc / template points (1 or 9) / then per source: type\n"
    "(Sx, Sz, Bulk, Sxz), strength, frequency [Hz], x z [m]\n\n"
    "Record file: nrec / then per record: component (Vx, Vz, txx, tzz, txz),\n"
    "Trace x z t0[s] dt[ms] or Snapshot t0[s] dt[ms]\n";

static void check(int st, const char *where)
{
    if (st != SM_OK) {
        fprintf(stderr, "Error in %s: %s (%s)\n", where, sm_strerror(st), sm_last_error());
        exit(1);
    }
}

static int recorded(int c)
{
    return c == SM_VX || c == SM_VZ || c == SM_TXX || c == SM_TZZ || c == SM_TXZ;
}

int main(int argc, char *argv[])
{
    sm_params p;
    sm_set_defaults(&p, 0);
    sm_process_args(argc, argv, &p, usage_text);
    const size_t np = p.np;   /* 0 when abs is 0 */
    const double dt = p.dt;
    if (sm_verbose) {
        fprintf(stdout, "\n *** ve_vti - ViscoElastic wave propagation in 2D VTI anisotropic media ***\n\n");
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
    int L = 0;
    sm_read_grid_ve(p.modelfile, &g, &L);
    const size_t nx2 = g.nx + 2 * np, nz2 = g.nz + 2 * np, n = nx2 * nz2;
    double *K = calloc(n, sizeof(double)), *mu = calloc(n, sizeof(double)), *rho = calloc(n, sizeof(double)),
           *eps = calloc(n, sizeof(double)), *delta = calloc(n, sizeof(double)),
           *Q1 = calloc((size_t)L * n, sizeof(double)), *f1 = calloc((size_t)L * n, sizeof(double)),
           *Q2 = calloc(n, sizeof(double)), *f2 = calloc(n, sizeof(double));
    if (sm_verbose) { fprintf(stdout, "done.\nReading properties of materials ... "); fflush(stdout); }
    sm_read_model_ve(p.modelfile, &g, np, L, K, mu, rho, eps, delta, Q1, f1, Q2, f2);

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
    sm_ve_params vp;
    memset(&vp, 0, sizeof(vp));
    vp.nx = g.nx; vp.nz = g.nz; vp.np = np; vp.dx = g.dx; vp.dz = g.dz; vp.dt = dt;
    vp.vmax = p.vmax; vp.pml_order = p.pml_order; vp.Rc = p.Rc; vp.kappa_max = p.kappa_max;
    vp.alpha_max = p.alpha_max; vp.f0 = ns ? src.s[0].f : 0.0; vp.abs = p.iwipe; vp.L = L;
    vp.storage = (stor && !strcmp(stor, "f32")) ? SM_STORE_F32 : SM_STORE_F64;
    if (sm_verbose) { fprintf(stdout, "done.\nPreprocessing data ... "); fflush(stdout); }
    /* the moduli files of ve_vti_pml.c:695-699 */
    {
        double *mod[5];
        for (int q = 0; q < 5; ++q) mod[q] = malloc(n * sizeof(double));
        check(sm_ve_moduli(nx2, nz2, L, K, mu, rho, eps, delta, Q1, f1, Q2, f2, mod[0], mod[1], mod[2], mod[3],
                           mod[4]), "sm_ve_moduli");
        const char *names[5] = {"c11", "c13", "c33", "c55", "K0"};
        for (int q = 0; q < 5; ++q) {
            sm_write_field_nc(mod[q], names[q], "MPa", g.nx, g.nz, np, g.x0, g.z0, g.dx, g.dz, out.basename, 1);
            free(mod[q]);
        }
    }
    sm_ve *sim = NULL;
    check(sm_ve_create(ctx, &vp, K, mu, rho, eps, delta, Q1, f1, Q2, f2, &sim), "sm_ve_create");
    check(sm_ve_set_sources(sim, ns, stype, sidx, len, fct), "sm_ve_set_sources");
    if (sm_verbose) fprintf(stdout, "done.\n");

    const size_t nsteps = sm_nsteps(p.duration, dt);   /* ve_vti_pml.c:714-715 */
    /* records of other components write nothing (ve_vti_pml.c:786-812) */
    size_t ntr = 0;
    for (size_t r = 0; r < out.nrec; ++r) ntr += out.r[r].type == REC_TRACE && recorded(out.r[r].comp);
    size_t *ridx = malloc((ntr ? ntr : 1) * sizeof(size_t));
    int *rcomp = malloc((ntr ? ntr : 1) * sizeof(int));
    size_t *rmap = malloc((ntr ? ntr : 1) * sizeof(size_t));
    for (size_t r = 0, q = 0; r < out.nrec; ++r)
        if (out.r[r].type == REC_TRACE && recorded(out.r[r].comp)) {
            ridx[q] = out.r[r].idx; rcomp[q] = out.r[r].comp; rmap[q] = r; ++q;
            if (p.segy == 1) {   /* ve_vti_pml.c:442-449 */
                out.r[r].nsamples = (size_t)round(p.duration / out.r[r].dt);
                out.r[r].data = calloc(out.r[r].nsamples + 1, sizeof(float));
            }
        }
    check(sm_ve_set_receivers(sim, ntr, ridx, rcomp, nsteps), "sm_ve_set_receivers");

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
        check(sm_ve_run(sim, it + 1 - done), "sm_ve_run");
        done = it + 1;
        for (size_t r = 0; r < out.nrec; ++r)
            if (out.r[r].type == REC_SNAPSHOT && recorded(out.r[r].comp) &&
                sm_record_due(it, dt, out.r[r].t0, out.r[r].dt)) {
                if (sm_verbose > 1) printf("Writing snapshot, t = %g s\n", t);
                check(sm_ve_get_field(sim, out.r[r].comp, field), "sm_ve_get_field");
                sm_write_snapshot_nc(field, t, g.nx, g.nz, np, g.x0, g.z0, g.dx, g.dz, out.basename, r,
                                     p.plot_strips);
            }
    }
    if (ntr) {
        double *tr = malloc(nsteps * ntr * sizeof(double));
        check(sm_ve_get_traces(sim, 0, nsteps, tr), "sm_ve_get_traces");
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
    check(sm_ve_destroy(sim), "sm_ve_destroy");
    check(sm_ctx_destroy(ctx), "sm_ctx_destroy");
    free(K); free(mu); free(rho); free(eps); free(delta); free(Q1); free(f1); free(Q2); free(f2);
    free(A); free(f); free(len); free(sidx); free(stype); free(fct); free(ridx); free(rcomp); free(rmap);
    free(field);
    sm_free_sources(&src);
    sm_free_records(&out);
    if (sm_verbose) fprintf(stdout, "\nEnd of computation\n");
    return 0;
}
