/*
 * a_iso_3d -- 3D acoustic wave propagation in isotropic media (staggered
 * grid, CPML), running the time loop on an MI355X through libseismod_hip.
 *
 * Drop-in for seismod's a_iso_3d: same command line, .par keywords, model,
 * source and record files, and the same trc/, ssh/ and segy/ outputs
 * (a_iso_3d.c:224-804).  Set-up and update arithmetic live in the library
 * (sm_a3d_*); this program reads the inputs, drives the time loop and
 * writes the records.
 *
 * Environment: SEISMOD_STORAGE = f32 (default), f64 (reference operation
 * order), bf16 or f16 (16-bit wavefield storage); SEISMOD_DEVICE selects the GPU.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "seismod_hip.h"
#include "sm_io.h"

static const char *usage_text =
    "Format of parameter file:\n\n"
    "value     # keyword, optional comment\n\n"
    "Keywords: model, source, output, basename, time (s), dt (ms), segy, shotpoint,\n"
    "abs, nl, plstr, kappa_max, alpha_max, Rc\n\n"
    "Model file: nx ny nz / dx dy dz / x0 y0 z0 / isotropic_acoustic_3d / then one\n"
    "\"Vp rho\" line per node, x and y outer, z inner (one column only: 1-D model)\n\n"
    "Source file: nsrc / 1 / then per source: type (Sx, Sz, Sxz, Sf, Bulk),\n"
    "strength, frequency [Hz], x y z [m]\n\n"
    "Record file: nrec / then per record: component (Vx, Vy, Vz, P),\n"
    "Trace x y z t0[s] dt[ms] or Snapshot t0[s] dt[ms]\n";

static void check(int st, const char *where)
{
    if (st != SM_OK) {
        fprintf(stderr, "Error in %s: %s (%s)\n", where, sm_strerror(st), sm_last_error());
        if (st == SM_ERR_CFL) {
            fprintf(stderr, "\n *** Warning!  Time step too large\naborting ***\n");
            abort();   /* a_iso_3d.c:411-417 */
        }
        exit(1);
    }
}

static int storage_from_env(void)
{
    const char *s = getenv("SEISMOD_STORAGE");
    if (!s || !strcmp(s, "f32")) return SM_STORE_F32;
    if (!strcmp(s, "f64")) return SM_STORE_F64;
    if (!strcmp(s, "bf16")) return SM_STORE_BF16;
    if (!strcmp(s, "f16")) return SM_STORE_F16;
    fprintf(stderr, "Error: SEISMOD_STORAGE must be f32, f64, bf16 or f16\n");
    exit(1);
}

int main(int argc, char *argv[])
{
    sm_params p;
    sm_set_defaults(&p, 1);
    sm_process_args(argc, argv, &p, usage_text);
    const double dt = p.dt;
    if (sm_verbose) {
        fprintf(stdout, "\n *** a_iso_3d - Acoustic wave propagation in 3D isotropic media ***\n\n");
        fprintf(stdout, "Reading size of model ... ");
    }
    sm_grid g;
    sm_read_grid3d(p.modelfile, &g);
    if (g.dx != g.dz) {
        fprintf(stderr, "Error: grid cell size must be equal in X and Z, aborting.\n");
        abort();   /* a_iso_3d.c:298-301 */
    }
    const size_t np = p.np;
    const size_t nx2 = g.nx + 2 * np, ny2 = g.ny + 2 * np, nz2 = g.nz + 2 * np;
    const size_t nnn = nx2 * ny2 * nz2;
    double *vp = calloc(nnn, sizeof(double)), *rho = calloc(nnn, sizeof(double));
    if (!vp || !rho) { fprintf(stderr, "Error: cannot allocate memory\n"); abort(); }
    if (sm_verbose) { fprintf(stdout, "done.\nReading properties of materials ... "); fflush(stdout); }
    sm_read_model_a(p.modelfile, &g, np, vp, rho);

    if (sm_verbose) { fprintf(stdout, "\nReading source parameters ... "); fflush(stdout); }
    sm_sources src;
    sm_read_source3d(p.sourcefile, &g, np, &src);
    if (src.nsrc == 0) { fprintf(stderr, "Error: no source in %s\n", p.sourcefile); return 1; }

    const char *dev = getenv("SEISMOD_DEVICE");
    sm_ctx *ctx = NULL;
    check(sm_ctx_create(dev ? atoi(dev) : 0, &ctx), "sm_ctx_create");
    sm_a3d_params ap;
    memset(&ap, 0, sizeof(ap));
    ap.nx = g.nx; ap.ny = g.ny; ap.nz = g.nz; ap.np = np;
    ap.dx = g.dx; ap.dy = g.dy; ap.dz = g.dz; ap.dt = dt;
    ap.vmax = p.vmax; ap.pml_order = p.pml_order; ap.Rc = p.Rc; ap.kappa_max = p.kappa_max;
    ap.alpha_max = p.alpha_max;
    ap.f0 = src.s[0].f;   /* a_iso_3d.c:426, 526 */
    ap.storage = storage_from_env();
    sm_a3d *sim = NULL;
    check(sm_a3d_create(ctx, &ap, vp, rho, &sim), "sm_a3d_create");

    const size_t ns = src.nsrc;
    double *A = malloc(ns * sizeof(double)), *f = malloc(ns * sizeof(double));
    size_t *len = malloc(ns * sizeof(size_t)), *sidx = malloc(ns * sizeof(size_t));
    size_t total = 0;
    for (size_t k = 0; k < ns; ++k) {
        A[k] = src.s[k].A; f[k] = src.s[k].f;
        len[k] = sm_ricker_length(f[k], dt);
        sidx[k] = src.s[k].idx;
        total += len[k];
    }
    double *fct = malloc(total * sizeof(double));
    check(sm_compute_src_fct(ns, A, f, dt, len, fct), "sm_compute_src_fct");
    check(sm_a3d_set_sources(sim, ns, sidx, len, fct), "sm_a3d_set_sources");

    if (sm_verbose) { fprintf(stdout, "done.\nReading output parameters ... "); fflush(stdout); }
    sm_records out;
    memset(&out, 0, sizeof(out));
    snprintf(out.basename, SM_NAME_LEN, "%s", p.basename);
    sm_read_output3d(p.outputfile, &g, np, &out);
    sm_check_dt_trc(&out, dt);
    if (sm_verbose) fprintf(stdout, "done.\n");

    const size_t nsteps = sm_nsteps(p.duration, dt);   /* a_iso_3d.c:578-579 */
    size_t ntr = 0;
    for (size_t r = 0; r < out.nrec; ++r) ntr += out.r[r].type == REC_TRACE;
    size_t *ridx = malloc((ntr ? ntr : 1) * sizeof(size_t));
    int *rcomp = malloc((ntr ? ntr : 1) * sizeof(int));
    size_t *rmap = malloc((ntr ? ntr : 1) * sizeof(size_t));
    for (size_t r = 0, q = 0; r < out.nrec; ++r)
        if (out.r[r].type == REC_TRACE) {
            ridx[q] = out.r[r].idx; rcomp[q] = out.r[r].comp; rmap[q] = r; ++q;
            if (p.segy == 1) {   /* a_iso_3d.c:454-462 */
                out.r[r].nsamples = (size_t)round(p.duration / out.r[r].dt);
                out.r[r].data = calloc(out.r[r].nsamples + 1, sizeof(float));
            }
        }
    check(sm_a3d_set_receivers(sim, ntr, ridx, rcomp, nsteps), "sm_a3d_set_receivers");

    if (sm_verbose) {
        fprintf(stdout, "Starting main loop (%zu iterations)\n", nsteps);
        fflush(stdout);
    }
    /* snapshot plane: iy = (y_src - y0)/dy, without +np (a_iso_3d.c:726) */
    const double fiy = (src.s[0].y - g.y0) / g.dy;
    const size_t iy = fiy > 0 ? (size_t)fiy : 0;
    double *slice = malloc(nx2 * nz2 * sizeof(double));
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
        check(sm_a3d_run(sim, it + 1 - done), "sm_a3d_run");
        done = it + 1;
        for (size_t r = 0; r < out.nrec; ++r)
            if (out.r[r].type == REC_SNAPSHOT && sm_record_due(it, dt, out.r[r].t0, out.r[r].dt)) {
                if (sm_verbose > 1) printf("Writing snapshot, t = %g s\n", t);
                check(sm_a3d_get_slice_xz(sim, out.r[r].comp, iy, slice), "sm_a3d_get_slice_xz");
                sm_write_snapshot_nc(slice, t, g.nx, g.nz, np, g.x0, g.z0, g.dx, g.dz, out.basename,
                                     r, p.plot_strips);
            }
    }
    if (ntr) {
        double *tr = malloc(nsteps * ntr * sizeof(double));
        check(sm_a3d_get_traces(sim, 0, nsteps, tr), "sm_a3d_get_traces");
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
    check(sm_a3d_destroy(sim), "sm_a3d_destroy");
    check(sm_ctx_destroy(ctx), "sm_ctx_destroy");
    free(vp); free(rho); free(A); free(f); free(len); free(sidx); free(fct);
    free(ridx); free(rcomp); free(rmap); free(slice);
    sm_free_sources(&src);
    sm_free_records(&out);
    if (sm_verbose) fprintf(stdout, "\nEnd of computation\n");
    return 0;
}
