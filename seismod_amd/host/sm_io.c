/*
 * sm_io.c -- input and output of the drop-in e_iso and a_iso_3d programs.
 *
 * Formats and behaviour follow seismod's io_utils.c (cited per function).
 * Where the reference reads or writes out of bounds or through an
 * uninitialised value, this code reports an error or uses the value the
 * reference evidently intends; each such place says so.
 */
#define _POSIX_C_SOURCE 200809L
#include "sm_io.h"

#include <ctype.h>
#include <errno.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>
#include <unistd.h>

int sm_verbose = 0;

static void die(const char *fmt, const char *arg)
{
    fprintf(stderr, fmt, arg);
    exit(1);
}

void sm_set_defaults(sm_params *p, int three_d)
{
    /* io_utils.c:63-135 */
    memset(p, 0, sizeof(*p));
    p->np = 20;
    p->vmax = 5000.0;
    p->pml_order = 2.0;
    p->Rc = 1.e-5;
    p->kappa_max = 1.0;
    p->alpha_max = 1;
    p->dt = 1.e-4;
    p->duration = 1.2;
    strcpy(p->basename, "run01");
    p->shotpt_no = 1;
    (void)three_d;
}

static void usage(const char *prog, FILE *out, int code, const char *text)
{
    /* print_usage, io_utils.c:137-154 */
    fprintf(out, "Usage: %s [options] -p parameter_file.dat\n", prog);
    fprintf(out, "\noptions are:\n");
    fprintf(out, "  -h    Print this message\n");
    fprintf(out, "  -v    Verbose mode (type twice for increased verbosity)\n\n");
    if (text) fprintf(out, "%s", text);
    fflush(out);
    exit(code);
}

void sm_process_args(int argc, char *const argv[], sm_params *p, const char *usage_text)
{
    /* io_utils.c:156-294 */
    char paramfile[256] = "";
    int have_p = 0, opt;
    while ((opt = getopt(argc, argv, "c:hp:v")) != -1) {
        switch (opt) {
        case 'c':
            /* io_utils.c:167-171 returns before the .par is read, so e_iso and
             * a_iso_3d would run on uninitialised file names; refuse instead */
            fprintf(stderr, "Error: checkpoint restart (-c) is not available for %s\n", argv[0]);
            exit(1);
        case 'h':
            usage(argv[0], stdout, 0, usage_text);
            break;
        case 'p':
            snprintf(paramfile, sizeof(paramfile), "%s", optarg);
            have_p = 1;
            break;
        case 'v':
            sm_verbose++;
            break;
        default:
            usage(argv[0], stderr, 1, usage_text);
        }
    }
    if (!have_p) usage(argv[0], stderr, 1, usage_text);

    FILE *fid = fopen(paramfile, "r");
    if (!fid) die("Error, cannot open parameter file %s for reading\n", paramfile);
    char line[200], value[SM_NAME_LEN], keyword[64];
    while (fgets(line, sizeof(line), fid)) {
        if (sscanf(line, "%79s", value) <= 0) continue;
        /* value # keyword, comment  (io_utils.c:214-222) */
        char *s1 = strchr(line, '#');
        if (!s1) continue;   /* the reference dereferences NULL here */
        s1++;
        char *s2 = strchr(s1, ',');
        if (s2) *s2 = '\0';
        if (sscanf(s1, "%63s", keyword) <= 0) continue;
        if (!strcmp(keyword, "model")) snprintf(p->modelfile, SM_NAME_LEN, "%s", value);
        else if (!strcmp(keyword, "source")) snprintf(p->sourcefile, SM_NAME_LEN, "%s", value);
        else if (!strcmp(keyword, "output")) snprintf(p->outputfile, SM_NAME_LEN, "%s", value);
        else if (!strcmp(keyword, "basename")) snprintf(p->basename, SM_NAME_LEN, "%s", value);
        else if (!strcmp(keyword, "time")) p->duration = strtod(value, NULL);
        else if (!strcmp(keyword, "dt")) p->dt = strtod(value, NULL) * 0.001;   /* ms -> s */
        else if (!strcmp(keyword, "abs")) p->iwipe = (short)atoi(value);
        else if (!strcmp(keyword, "nl")) p->np = (size_t)atoi(value);
        else if (!strcmp(keyword, "plstr")) p->plot_strips = (short)atoi(value);
        else if (!strcmp(keyword, "segy")) p->segy = (short)atoi(value);
        else if (!strcmp(keyword, "Rc")) p->Rc = strtod(value, NULL);
        else if (!strcmp(keyword, "kappa_max")) p->kappa_max = strtod(value, NULL);
        else if (!strcmp(keyword, "alpha_max")) p->alpha_max = (short)atoi(value);
        else if (!strcmp(keyword, "shotpoint")) p->shotpt_no = atoi(value);
        else if (!strcmp(keyword, "simulateCMP")) p->simulate_cmp = atoi(value);
        /* other keywords (check_model, chkpt_inc, pml_median, saveEnergy,
         * EnergyROI, azimuthal_mode) do not affect these programs */
    }
    fclose(fid);
    if (p->iwipe == 0) p->np = 0;   /* io_utils.c:293 */
}

static FILE *open_or_die(const char *fn)
{
    FILE *f = fopen(fn, "r");
    if (!f) die("Error, cannot open %s for reading\n", fn);
    return f;
}

void sm_read_grid2d(const char *fn, sm_grid *g)
{
    /* io_utils.c:296-328 (e_iso calls it with np = 0, so nx2 = nx) */
    FILE *f = open_or_die(fn);
    memset(g, 0, sizeof(*g));
    if (fscanf(f, "%zu %zu", &g->nx, &g->nz) != 2 || fscanf(f, "%lf %lf", &g->dx, &g->dz) != 2 ||
        fscanf(f, "%lf %lf", &g->x0, &g->z0) != 2)
        die("Error: something wrong with file %s\n", fn);
    g->ny = 1;
    g->dy = g->dx;
    fclose(f);
}

void sm_read_grid3d(const char *fn, sm_grid *g)
{
    /* io_utils.c:384-401 */
    FILE *f = open_or_die(fn);
    memset(g, 0, sizeof(*g));
    if (fscanf(f, "%zu %zu %zu", &g->nx, &g->ny, &g->nz) != 3 ||
        fscanf(f, "%lf %lf %lf", &g->dx, &g->dy, &g->dz) != 3 ||
        fscanf(f, "%lf %lf %lf", &g->x0, &g->y0, &g->z0) != 3)
        die("Error: something wrong with file %s\n", fn);
    fclose(f);
}

void sm_read_model_e(const char *fn, const sm_grid *g, double *vp, double *vs, double *rho)
{
    /* io_utils.c:892-947 */
    FILE *f = open_or_die(fn);
    char type[64];
    if (fscanf(f, "%*d %*d") < 0 || fscanf(f, "%*f %*f") < 0 || fscanf(f, "%*f %*f") < 0 ||
        fscanf(f, "%63s", type) != 1 || strcmp(type, "isotropic_elastic") != 0)
        die("Error, model in %s not isotropic_elastic\n", fn);
    int flag1d = 0;
    for (size_t i = 0; i < g->nx; ++i) {
        size_t ind = i * g->nz;
        int nread = fscanf(f, "%lf %lf %lf", &vp[ind], &vs[ind], &rho[ind]);
        if (nread <= 0 && i == 1) { flag1d = 1; break; }   /* one column given */
        for (size_t j = 1; j < g->nz; ++j) {
            ind = i * g->nz + j;
            if (fscanf(f, "%lf %lf %lf", &vp[ind], &vs[ind], &rho[ind]) != 3)
                die("Error: something wrong with file %s\n", fn);
        }
    }
    fclose(f);
    if (flag1d)
        for (size_t i = 1; i < g->nx; ++i)
            for (size_t j = 0; j < g->nz; ++j) {
                vp[i * g->nz + j] = vp[j];
                vs[i * g->nz + j] = vs[j];
                rho[i * g->nz + j] = rho[j];
            }
}

void sm_read_model_a(const char *fn, const sm_grid *g, size_t np, double *vp, double *rho)
{
    /* io_utils.c:1029-1153 */
    FILE *f = open_or_die(fn);
    char type[64];
    if (fscanf(f, "%*d %*d %*d") < 0 || fscanf(f, "%*f %*f %*f") < 0 ||
        fscanf(f, "%*f %*f %*f") < 0 || fscanf(f, "%63s", type) != 1 ||
        strcmp(type, "isotropic_acoustic_3d") != 0)
        die("Error, model in %s not isotropic_acoustic_3d\n", fn);
    const size_t ny2 = g->ny + 2 * np, nz2 = g->nz + 2 * np, nx2 = g->nx + 2 * np;
#define AT(i, j, k) ((((i)) * ny2 + (j)) * nz2 + (k))
    int flag1d = 0;
    for (size_t i = 0; i < g->nx && !flag1d; ++i)
        for (size_t j = 0; j < g->ny; ++j) {
            size_t ind = AT(i + np, j + np, np);
            int nread = fscanf(f, "%lf %lf", &vp[ind], &rho[ind]);
            if (nread <= 0 && j == 1) { flag1d = 1; break; }   /* one column given */
            for (size_t k = 1; k < g->nz; ++k) {
                ind = AT(i + np, j + np, k + np);
                if (fscanf(f, "%lf %lf", &vp[ind], &rho[ind]) != 2)
                    die("Error: something wrong with file %s\n", fn);
            }
        }
    fclose(f);
    if (flag1d)
        for (size_t i = 0; i < g->nx; ++i)
            for (size_t j = 0; j < g->ny; ++j)
                for (size_t k = 0; k < g->nz; ++k) {
                    vp[AT(i + np, j + np, k + np)] = vp[AT(np, np, k + np)];
                    rho[AT(i + np, j + np, k + np)] = rho[AT(np, np, k + np)];
                }
    /* edge values into the absorbing layers: z, then y, then x */
    for (size_t i = 0; i < g->nx; ++i)
        for (size_t j = 0; j < g->ny; ++j)
            for (size_t k = 0; k < np; ++k) {
                vp[AT(i + np, j + np, k)] = vp[AT(i + np, j + np, np)];
                rho[AT(i + np, j + np, k)] = rho[AT(i + np, j + np, np)];
                vp[AT(i + np, j + np, np + g->nz + k)] = vp[AT(i + np, j + np, np + g->nz - 1)];
                rho[AT(i + np, j + np, np + g->nz + k)] = rho[AT(i + np, j + np, np + g->nz - 1)];
            }
    for (size_t i = 0; i < g->nx; ++i)
        for (size_t k = 0; k < nz2; ++k)
            for (size_t j = 0; j < np; ++j) {
                vp[AT(i + np, j, k)] = vp[AT(i + np, np, k)];
                rho[AT(i + np, j, k)] = rho[AT(i + np, np, k)];
                vp[AT(i + np, np + g->ny + j, k)] = vp[AT(i + np, np + g->ny - 1, k)];
                rho[AT(i + np, np + g->ny + j, k)] = rho[AT(i + np, np + g->ny - 1, k)];
            }
    for (size_t j = 0; j < ny2; ++j)
        for (size_t k = 0; k < nz2; ++k)
            for (size_t i = 0; i < np; ++i) {
                vp[AT(i, j, k)] = vp[AT(np, j, k)];
                rho[AT(i, j, k)] = rho[AT(np, j, k)];
                vp[AT(np + g->nx + i, j, k)] = vp[AT(np + g->nx - 1, j, k)];
                rho[AT(np + g->nx + i, j, k)] = rho[AT(np + g->nx - 1, j, k)];
            }
    (void)nx2;
#undef AT
}

static int source_type(const char *t, int three_d)
{
    /* io_utils.c:1451-1481 (2D), 1561-1575 (3D); enum sourceType numbering */
    struct { const char *a, *b, *c; int v; int in3d; } tab[] = {
        {"sx", "SX", "Sx", 0, 1}, {"sz", "SZ", "Sz", 2, 1}, {"sxy", "SXY", "Sxy", 3, 0},
        {"sxz", "SXZ", "Sxz", 4, 1}, {"srz", "SRZ", "Srz", 4, 0}, {"syz", "SYZ", "Syz", 5, 0},
        {"sf", "SF", "Sf", 7, 1}, {"bulk", "BULK", "Bulk", 6, 1}, {"bulk_s", "BULK_S", "Bulk_s", 8, 0},
        {"fr", "FR", "Fr", 9, 0}, {"fz", "FZ", "Fz", 11, 0}, {"ftheta", "FTHETA", "Ftheta", 10, 0},
        {"kurkjian", "Kurkjian", "kurkjian", 12, 0}};
    for (size_t n = 0; n < sizeof(tab) / sizeof(tab[0]); ++n)
        if ((!three_d || tab[n].in3d) &&
            (!strcmp(t, tab[n].a) || !strcmp(t, tab[n].b) || !strcmp(t, tab[n].c)))
            return tab[n].v;
    return -1;
}

/* grid index of a coordinate, round() as the reference; out of range -> error */
static size_t node(double x, double x0, double d, size_t n, const char *what)
{
    const double r = round((x - x0) / d);
    if (!(r >= 0.0) || r >= (double)n) {
        fprintf(stderr, "Error: %s coordinate %g outside the grid\n", what, x);
        exit(1);
    }
    return (size_t)r;
}

void sm_read_source2d(const char *fn, const sm_grid *g, sm_sources *src)
{
    /* io_utils.c:1432-1538, with np = 0 as e_iso uses it */
    FILE *f = open_or_die(fn);
    memset(src, 0, sizeof(*src));
    if (fscanf(f, "%zu", &src->nsrc) != 1) die("Error: something wrong with file %s\n", fn);
    if (fscanf(f, "%zu", &src->ntemplate) != 1) src->ntemplate = 1;
    if (src->ntemplate != 1 && src->ntemplate != 9)
        die("Error: source template must have 1 or 9 points (%s)\n", fn);
    src->s = calloc(src->nsrc * src->ntemplate, sizeof(sm_source));
    char t[64];
    for (size_t n = 0; n < src->nsrc * src->ntemplate; n += src->ntemplate) {
        sm_source *s = &src->s[n];
        if (fscanf(f, "%63s", t) != 1) die("Error: something wrong with file %s\n", fn);
        s->type = source_type(t, 0);
        if (s->type < 0) die("Error: type of source (%s) not implemented\n", t);
        if (fscanf(f, "%lf", &s->A) != 1 || fscanf(f, "%lf", &s->f) != 1 ||
            fscanf(f, "%lf %lf", &s->x, &s->z) != 2)
            die("Error: something wrong with file %s\n", fn);
        size_t i = node(s->x, g->x0, g->dx, g->nx, "source x");
        size_t j = node(s->z, g->z0, g->dz, g->nz, "source z");
        s->idx = i * g->nz + j;
        if (src->ntemplate == 9) {
            /* 9-point template (io_utils.c:1494-1535); the reference leaves the
             * type of the extra points unset, here they take the main type */
            static const double w[9] = {0.5, 0.125, 0.25, 0.125, 0.25, 0.25, 0.125, 0.25, 0.125};
            static const int di[9] = {0, -1, 0, 1, -1, 1, -1, 0, 1};
            static const int dj[9] = {0, -1, -1, -1, 0, 0, 1, 1, 1};
            const double A = s->A;
            for (int q = 0; q < 9; ++q) {
                sm_source *e = &src->s[n + q];
                e->type = s->type;
                e->A = w[q] * A;
                e->f = s->f;
                e->x = s->x; e->z = s->z;
                size_t ii = node(s->x + di[q] * g->dx, g->x0, g->dx, g->nx, "template x");
                size_t jj = node(s->z + dj[q] * g->dz, g->z0, g->dz, g->nz, "template z");
                e->idx = ii * g->nz + jj;
            }
        }
    }
    fclose(f);
}

void sm_read_source3d(const char *fn, const sm_grid *g, size_t np, sm_sources *src)
{
    /* io_utils.c:1540-1589 */
    FILE *f = open_or_die(fn);
    memset(src, 0, sizeof(*src));
    if (fscanf(f, "%zu", &src->nsrc) != 1 || fscanf(f, "%zu", &src->ntemplate) != 1)
        die("Error: something wrong with file %s\n", fn);
    if (src->ntemplate != 1) {
        fprintf(stdout, "Warning, source template parameter ignored.\n");
        src->ntemplate = 1;
    }
    src->s = calloc(src->nsrc, sizeof(sm_source));
    const size_t ny2 = g->ny + 2 * np, nz2 = g->nz + 2 * np;
    char t[64];
    for (size_t n = 0; n < src->nsrc; ++n) {
        sm_source *s = &src->s[n];
        if (fscanf(f, "%63s", t) != 1) die("Error: something wrong with file %s\n", fn);
        s->type = source_type(t, 1);
        if (s->type < 0) die("Error: type of source (%s) not implemented\n", t);
        if (fscanf(f, "%lf", &s->A) != 1 || fscanf(f, "%lf", &s->f) != 1 ||
            fscanf(f, "%lf %lf %lf", &s->x, &s->y, &s->z) != 3)
            die("Error: something wrong with file %s\n", fn);
        size_t i = node(s->x, g->x0, g->dx, g->nx, "source x") + np;
        size_t j = node(s->y, g->y0, g->dy, g->ny, "source y") + np;
        size_t k = node(s->z, g->z0, g->dz, g->nz, "source z") + np;
        s->idx = (i * ny2 + j) * nz2 + k;
    }
    fclose(f);
}

static int comp_code(const char *t, int three_d)
{
    /* io_utils.c:1607-1651 (2D), 1700-1711 (3D); enum component numbering */
    struct { const char *a, *b, *c; int v; int in3d; } tab[] = {
        {"Vx", "vx", "VX", 0, 1}, {"Vy", "vy", "VY", 1, 1}, {"Vz", "vz", "VZ", 2, 1},
        {"Qx", "qx", "QX", 3, 0}, {"Qz", "qz", "QZ", 4, 0}, {"txx", "TXX", "TXX", 5, 0},
        {"tzz", "TZZ", "TZZ", 6, 0}, {"txy", "TXY", "TXY", 9, 0}, {"tyz", "TYZ", "TYZ", 10, 0},
        {"txz", "TXZ", "TXZ", 7, 0}, {"Vr", "vr", "VR", 13, 0}, {"Vtheta", "vtheta", "VTHETA", 14, 0},
        {"trr", "TRR", "TRR", 15, 0}, {"trtheta", "TRTHETA", "TRTHETA", 16, 0},
        {"trz", "TRZ", "TRZ", 17, 0}, {"tthetatheta", "TTHETATHETA", "TTHETATHETA", 18, 0},
        {"tthetaz", "TTHETAZ", "TTHETAZ", 19, 0}, {"p", "P", "P", 8, 1},
        {"div", "Div", "DIV", 11, 0}, {"curl", "Curl", "CURL", 12, 0}};
    for (size_t n = 0; n < sizeof(tab) / sizeof(tab[0]); ++n)
        if ((!three_d || tab[n].in3d) &&
            (!strcmp(t, tab[n].a) || !strcmp(t, tab[n].b) || !strcmp(t, tab[n].c)))
            return tab[n].v;
    return -1;
}

static void read_output(const char *fn, const sm_grid *g, size_t np, sm_records *out, int three_d)
{
    FILE *f = open_or_die(fn);
    if (fscanf(f, "%zu", &out->nrec) != 1) die("Error: something wrong with file %s\n", fn);
    out->r = calloc(out->nrec ? out->nrec : 1, sizeof(sm_record));
    const size_t ny2 = g->ny + 2 * np, nz2 = g->nz + 2 * np;
    char t[80];
    for (size_t n = 0; n < out->nrec; ++n) {
        sm_record *r = &out->r[n];
        if (fscanf(f, "%79s", t) != 1) die("Error: something wrong with file %s\n", fn);
        r->comp = comp_code(t, three_d);
        if (r->comp < 0) die("Error: particle velocity component (%s) does not exists\n", t);
        if (fscanf(f, "%79s", t) != 1) die("Error: something wrong with file %s\n", fn);
        if (!strcmp(t, "Trace") || !strcmp(t, "trace") || !strcmp(t, "TRACE")) {
            r->type = REC_TRACE;
            int ok = three_d ? fscanf(f, "%lf %lf %lf", &r->x, &r->y, &r->z) == 3
                             : fscanf(f, "%lf %lf", &r->x, &r->z) == 2;
            if (!ok || fscanf(f, "%lf", &r->t0) != 1 || fscanf(f, "%lf", &r->dt) != 1)
                die("Error: something wrong with file %s\n", fn);
            r->dt *= 1.e-3;   /* ms -> s */
            /* lround, io_utils.c:1661-1666, 1722-1725 */
            const double fi = (double)lround((r->x - g->x0) / g->dx);
            const double fk = (double)lround((r->z - g->z0) / g->dz);
            const double fj = three_d ? (double)lround((r->y - g->y0) / g->dy) : 0.0;
            if (fi < 0 || fi >= g->nx || fk < 0 || fk >= g->nz || fj < 0 || fj >= g->ny) {
                fprintf(stderr, "Error: record %zu is outside the grid\n", n + 1);
                exit(1);
            }
            const size_t i = (size_t)fi + np, j = (size_t)fj + np, k = (size_t)fk + np;
            r->idx = three_d ? (i * ny2 + j) * nz2 + k : i * g->nz + k;
        } else if (!strcmp(t, "Snapshot") || !strcmp(t, "snapshot") || !strcmp(t, "SNAPSHOT")) {
            r->type = REC_SNAPSHOT;
            if (fscanf(f, "%lf", &r->t0) != 1 || fscanf(f, "%lf", &r->dt) != 1)
                die("Error: something wrong with file %s\n", fn);
            r->dt *= 1.e-3;
        } else {
            die("Error: type of record (%s) not implemented\n", t);
        }
        r->fid = NULL;
    }
    fclose(f);
}

void sm_read_output2d(const char *fn, const sm_grid *g, sm_records *out)
{
    read_output(fn, g, 0, out, 0);   /* io_utils.c:1592-1682 */
}

void sm_read_output3d(const char *fn, const sm_grid *g, size_t np, sm_records *out)
{
    read_output(fn, g, np, out, 1);  /* io_utils.c:1684-1741 */
}

void sm_check_dt_trc(const sm_records *out, double dt)
{
    /* io_utils.c:3561-3574 */
    for (size_t n = 0; n < out->nrec; ++n)
        if (out->r[n].type == REC_TRACE) {
            double rem = fmod(1000000.0 * out->r[n].dt, 1000000.0 * dt);
            if (rem > 1.e-9) {
                fprintf(stdout, "\n *** Warning ***\n");
                fprintf(stdout, "Time step of receiver %zu not a multiple of modeling time step\n", n + 1);
                fprintf(stdout, " *** Warning ***\n");
            }
        }
}

static void make_dir(const char *path)
{
    struct stat sb;
    if (stat(path, &sb) < 0 && errno == ENOENT) mkdir(path, S_IRWXU);
}

void sm_write_trace(sm_records *out, size_t nr, double t, double value)
{
    /* io_utils.c:1744-1768 */
    sm_record *r = &out->r[nr];
    if (!r->fid) {
        char fn[3 * SM_NAME_LEN + 32];
        make_dir("trc");
        snprintf(fn, sizeof(fn), "trc/%s", out->basename);
        make_dir(fn);
        snprintf(fn, sizeof(fn), "trc/%s/%s_%zu.trc", out->basename, out->basename, nr + 1);
        r->fid = fopen(fn, "w+");
        if (!r->fid) {
            fprintf(stderr, "Error, cannot open %s for output (%s)\n", fn, strerror(errno));
            exit(1);
        }
    }
    fprintf(r->fid, "%le    %le\n", t, 1.e3 * value);
}

void sm_close_traces(sm_records *out)
{
    for (size_t n = 0; n < out->nrec; ++n)
        if (out->r[n].fid) { fclose(out->r[n].fid); out->r[n].fid = NULL; }
}

/* ------------------------------------------------------------------ */
/* netCDF-3 classic (CDF-1) writer                                    */
/* ------------------------------------------------------------------ */
typedef struct {
    unsigned char *b;
    size_t n, cap;
} buf;

static void put(buf *o, const void *p, size_t n)
{
    if (o->n + n > o->cap) {
        o->cap = (o->n + n) * 2 + 256;
        o->b = realloc(o->b, o->cap);
    }
    memcpy(o->b + o->n, p, n);
    o->n += n;
}
static void put_u32(buf *o, uint32_t v)
{
    unsigned char c[4] = {(unsigned char)(v >> 24), (unsigned char)(v >> 16),
                          (unsigned char)(v >> 8), (unsigned char)v};
    put(o, c, 4);
}
static void put_pad(buf *o)
{
    static const unsigned char z[4] = {0, 0, 0, 0};
    put(o, z, (4 - o->n % 4) % 4);
}
static void put_name(buf *o, const char *s)
{
    put_u32(o, (uint32_t)strlen(s));
    put(o, s, strlen(s));
    put_pad(o);
}
static void put_f64(buf *o, double d)
{
    uint64_t u;
    memcpy(&u, &d, 8);
    put_u32(o, (uint32_t)(u >> 32));
    put_u32(o, (uint32_t)u);
}
static void put_f32(buf *o, float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    put_u32(o, u);
}

enum { NC_CHAR = 2, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6 };
enum { TAG_DIM = 10, TAG_VAR = 11, TAG_ATT = 12 };

/* text attribute of len bytes (the reference's lengths count the NUL) */
static void att_text(buf *o, const char *name, const char *text, size_t len)
{
    put_name(o, name);
    put_u32(o, NC_CHAR);
    put_u32(o, (uint32_t)len);
    for (size_t i = 0; i < len; ++i) {
        char c = i < strlen(text) ? text[i] : '\0';
        put(o, &c, 1);
    }
    put_pad(o);
}
static void att_f64(buf *o, const char *name, const double *v, size_t n)
{
    put_name(o, name);
    put_u32(o, NC_DOUBLE);
    put_u32(o, (uint32_t)n);
    for (size_t i = 0; i < n; ++i) put_f64(o, v[i]);
}

int sm_write_snapshot_nc(const double *data, double t, size_t nx, size_t nz, size_t np,
                         double x0, double z0, double dx, double dz, const char *basename,
                         size_t nr, int plstr)
{
    /* write_snapshot_nc / write_snapshot3d_nc (io_utils.c:1822-2180): dims
     * x, y; double coordinate variables x(x), y(y); float z(y, x) in mm/s */
    char fn[3 * SM_NAME_LEN + 64];
    make_dir("ssh");
    snprintf(fn, sizeof(fn), "ssh/%s", basename);
    make_dir(fn);
    snprintf(fn, sizeof(fn), "ssh/%s/%s_%zu_%lf.nc", basename, basename, nr + 1, t);
    const size_t nx2 = nx + 2 * np, nz2 = nz + 2 * np;
    const size_t mx = plstr ? nx2 : nx, mz = plstr ? nz2 : nz;
    const size_t ox = plstr ? 0 : np, oz = plstr ? 0 : np;
    double xr[2], yr[2], zr[2];
    if (plstr) {
        xr[0] = x0 - np * dx; xr[1] = x0 + (nx + np - 1) * dx;
        yr[0] = z0 - np * dz; yr[1] = z0 + (nz + np - 1) * dz;
    } else {
        xr[0] = x0; xr[1] = x0 + (nx - 1) * dx;
        yr[0] = z0; yr[1] = z0 + (nz - 1) * dz;
    }
    zr[0] = zr[1] = 1.e3 * data[ox * nz2 + oz];
    for (size_t i = 0; i < mx; ++i)
        for (size_t j = 0; j < mz; ++j) {
            const double v = 1.e3 * data[(i + ox) * nz2 + j + oz];
            zr[0] = zr[0] < v ? zr[0] : v;
            zr[1] = zr[1] > v ? zr[1] : v;
        }
    buf h = {0};
    put(&h, "CDF\x01", 4);
    put_u32(&h, 0);                          /* numrecs */
    put_u32(&h, TAG_DIM); put_u32(&h, 2);     /* dims: x (id 0), y (id 1) */
    put_name(&h, "x"); put_u32(&h, (uint32_t)mx);
    put_name(&h, "y"); put_u32(&h, (uint32_t)mz);
    put_u32(&h, TAG_ATT); put_u32(&h, 3);     /* global attributes */
    att_text(&h, "Conventions", "COARDS/CF-1.0", 14);
    att_text(&h, "title", "pve_iso", 8);
    put_name(&h, "node_offset"); put_u32(&h, NC_INT); put_u32(&h, 1); put_u32(&h, 0);
    put_u32(&h, TAG_VAR); put_u32(&h, 3);
    /* variable records; their data offsets are patched once known */
    size_t begin_at[3];
    const uint32_t vsize[3] = {(uint32_t)(8 * mx), (uint32_t)(8 * mz), (uint32_t)(4 * mx * mz)};
    for (int v = 0; v < 3; ++v) {
        if (v == 0) { put_name(&h, "x"); put_u32(&h, 1); put_u32(&h, 0); }
        else if (v == 1) { put_name(&h, "y"); put_u32(&h, 1); put_u32(&h, 1); }
        else { put_name(&h, "z"); put_u32(&h, 2); put_u32(&h, 1); put_u32(&h, 0); }
        put_u32(&h, TAG_ATT);
        if (v < 2) {
            put_u32(&h, 3);
            att_text(&h, "long_name", v == 0 ? "X" : "Y", 2);
            att_text(&h, "units", "m", 2);
            att_f64(&h, "actual_range", v == 0 ? xr : yr, 2);
            put_u32(&h, NC_DOUBLE);
        } else {
            put_u32(&h, 4);
            att_text(&h, "units", "mm/s", 5);
            const double one = 1.0;
            att_f64(&h, "scale_factor", &one, 1);
            put_name(&h, "_FillValue"); put_u32(&h, NC_FLOAT); put_u32(&h, 1);
            put_f32(&h, (float)fabs(NAN));
            att_f64(&h, "actual_range", zr, 2);
            put_u32(&h, NC_FLOAT);
        }
        put_u32(&h, (vsize[v] + 3) / 4 * 4);
        begin_at[v] = h.n;
        put_u32(&h, 0);
    }
    uint32_t off = (uint32_t)h.n;
    for (int v = 0; v < 3; ++v) {
        unsigned char *p = h.b + begin_at[v];
        p[0] = off >> 24; p[1] = off >> 16; p[2] = off >> 8; p[3] = off;
        off += (vsize[v] + 3) / 4 * 4;
    }
    /* data: x, y coordinates then z row by row (y outer) */
    for (size_t i = 0; i < mx; ++i)
        put_f64(&h, plstr ? x0 + ((1.0 * i) - np) * dx : x0 + i * dx);
    for (size_t j = 0; j < mz; ++j)
        put_f64(&h, plstr ? z0 + ((1.0 * j) - np) * dz : z0 + j * dz);
    for (size_t j = 0; j < mz; ++j)
        for (size_t i = 0; i < mx; ++i)
            put_f32(&h, (float)data[(i + ox) * nz2 + j + oz] * 1.e3f);
    put_pad(&h);
    FILE *f = fopen(fn, "wb");
    if (!f) {
        fprintf(stderr, "Error, cannot open %s for output (%s)\n", fn, strerror(errno));
        free(h.b);
        return 1;
    }
    fwrite(h.b, 1, h.n, f);
    fclose(f);
    free(h.b);
    return 0;
}

/* ------------------------------------------------------------------ */
/* SEG-Y (io_utils.c:3269-3559), host byte order as the reference     */
/* ------------------------------------------------------------------ */
static void w16(FILE *f, int16_t v) { fwrite(&v, 2, 1, f); }
static void w32(FILE *f, int32_t v) { fwrite(&v, 4, 1, f); }

int sm_save_segy(const sm_params *p, const sm_records *out, const sm_sources *src)
{
    char fn[SM_NAME_LEN + 16];
    make_dir("segy");
    snprintf(fn, sizeof(fn), "segy/%s.sgy", p->basename);
    FILE *f = fopen(fn, "w");
    if (!f) {
        fprintf(stderr, "Error, cannot open %s for input (%s)\n", fn, strerror(errno));
        exit(1);
    }
    if (sm_verbose > 1) printf("Saving traces in file %s (SEG Y format) ... ", fn);
    char text[3200];
    memset(text, ' ', sizeof(text));
    text[0] = 'C';
    text[80] = 'C';
    fwrite(text, 1, 3200, f);
    w32(f, 1); w32(f, 1); w32(f, 1);            /* job, line, reel */
    int16_t ntr = 0;
    for (size_t n = 0; n < out->nrec; ++n) ntr += out->r[n].type == REC_TRACE;
    w16(f, ntr);
    w16(f, 0);
    int16_t si = 0, ns = 0;
    for (size_t n = 0; n < out->nrec; ++n)
        if (out->r[n].type == REC_TRACE) {
            si = (int16_t)(1000000 * out->r[n].dt);
            ns = (int16_t)round(p->duration / out->r[n].dt);
            break;   /* all traces are taken to share one interval */
        }
    w16(f, si); w16(f, si); w16(f, ns); w16(f, ns);
    w16(f, 5);                                  /* IEEE floats */
    int16_t fold = 0;
    for (size_t n = 0; n < out->nrec; ++n) fold += out->r[n].type == REC_TRACE && out->r[n].comp == 2;
    if (fold == 0)
        for (size_t n = 0; n < out->nrec; ++n) fold += out->r[n].type == REC_TRACE && out->r[n].comp == 1;
    w16(f, fold);
    w16(f, 1); w16(f, 1);
    for (int q = 0; q < 10; ++q) w16(f, 0);
    w16(f, 1); w16(f, 1); w16(f, 1);
    fseek(f, 3500L, SEEK_SET);
    uint16_t rev = 1;
    fwrite(&rev, 2, 1, f);
    w16(f, 0); w16(f, 0);
    fseek(f, 3600L, SEEK_SET);
    int32_t tr_no = 1, src_no = p->shotpt_no;
    const sm_source *s0 = &src->s[0];
    for (size_t n = 0; n < out->nrec; ++n) {
        const sm_record *r = &out->r[n];
        if (r->type != REC_TRACE) continue;
        if (r->comp != 2 && r->comp != 1 && r->comp != 0 && r->comp != 11 && r->comp != 12) continue;
        w32(f, tr_no); w32(f, tr_no); w32(f, tr_no); w32(f, tr_no);
        w32(f, src_no);
        if (p->simulate_cmp != 0) src_no++;
        w32(f, 1);
        w32(f, tr_no);
        tr_no++;
        int16_t code = r->comp == 2 ? 12 : r->comp == 1 ? 13 : r->comp == 0 ? 14 : r->comp == 11 ? 23 : 24;
        w16(f, code);
        w16(f, 1); w16(f, 1); w16(f, 1);
        const int32_t offset = (int32_t)(r->x - s0->x);
        w32(f, offset);
        w32(f, (int32_t)(-r->z));
        w32(f, 0);
        w32(f, (int32_t)(s0->z - 0));
        w32(f, 0); w32(f, 0); w32(f, 0); w32(f, 0);
        w16(f, 1); w16(f, 1);
        w32(f, p->simulate_cmp == 0 ? (int32_t)s0->x : -offset / 2);
        w32(f, 0);
        w32(f, p->simulate_cmp == 0 ? (int32_t)r->x : offset / 2);
        w32(f, 0);
        w16(f, 1); w16(f, 500); w16(f, 500);
        for (int q = 0; q < 10; ++q) w16(f, 0);
        const int16_t nsr = (int16_t)round(p->duration / r->dt);
        w16(f, nsr);
        w16(f, (int16_t)(1000000 * r->dt));
        w16(f, 1); w16(f, 1); w16(f, 1);
        for (int q = 0; q < 16; ++q) w16(f, 0);
        time_t now = time(NULL);
        struct tm *tm_now = gmtime(&now);
        w16(f, (int16_t)(1900 + tm_now->tm_year));
        w16(f, (int16_t)tm_now->tm_yday);
        w16(f, (int16_t)tm_now->tm_hour);
        w16(f, (int16_t)tm_now->tm_min);
        w16(f, (int16_t)tm_now->tm_sec);
        w16(f, 2);
        for (int q = 0; q < 5; ++q) w16(f, 0);
        w16(f, 2);
        w32(f, offset);
        w32(f, 0); w32(f, 0); w32(f, 0); w32(f, 0);
        w16(f, 1); w16(f, 6);
        w32(f, 1);
        w16(f, 0); w16(f, 6); w16(f, 0); w16(f, 1); w16(f, 4);
        fseek(f, 22L, SEEK_CUR);
        /* samples; the reference allocates round(duration/rdt) values and
         * indexes them by round(t/rdt), so its last sample lands one past the
         * end: here that sample is dropped and sample 0 stays 0 */
        for (int16_t q = 0; q < nsr; ++q) {
            float v = (r->data && (size_t)q < r->nsamples) ? r->data[q] : 0.f;
            fwrite(&v, 4, 1, f);
        }
    }
    fclose(f);
    if (sm_verbose > 1) printf("done.\n");
    return 0;
}

void sm_free_sources(sm_sources *s)
{
    free(s->s);
    s->s = NULL;
}

void sm_free_records(sm_records *r)
{
    for (size_t n = 0; n < r->nrec; ++n) free(r->r[n].data);
    free(r->r);
    r->r = NULL;
}
