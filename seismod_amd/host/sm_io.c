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

#include <complex.h>
#include <ctype.h>
#include <errno.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>
#include <unistd.h>
#ifdef _OPENMP
#include <omp.h>
#endif

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
    sm_process_args_ck(argc, argv, p, usage_text, NULL, 0);
}

int sm_process_args_ck(int argc, char *const argv[], sm_params *p, const char *usage_text, char *ckfile,
                       size_t ckfile_len)
{
    /* io_utils.c:156-294 */
    char paramfile[256] = "";
    int have_p = 0, have_c = 0, opt;
    while ((opt = getopt(argc, argv, "c:hp:v")) != -1) {
        switch (opt) {
        case 'c':
            /* io_utils.c:167-171 returns before the .par is read, so e_iso and
             * a_iso_3d would run on uninitialised file names; programs
             * without checkpoints refuse instead */
            if (!ckfile) {
                fprintf(stderr, "Error: checkpoint restart (-c) is not available for %s\n", argv[0]);
                exit(1);
            }
            snprintf(ckfile, ckfile_len, "%s", optarg);
            have_c = 1;
            break;
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
    if (have_c) return 1;   /* the parameters come from the checkpoint; -p is ignored */
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
        else if (!strcmp(keyword, "azimuthal_mode")) p->mode = atoi(value);
        else if (!strcmp(keyword, "check_model")) p->check_model = (short)atoi(value);
        else if (!strcmp(keyword, "chkpt_inc")) p->chkpt_inc = atoi(value);
        else if (!strcmp(keyword, "pml_median")) p->pml_median = (short)atoi(value);
        else if (!strcmp(keyword, "saveEnergy")) p->save_energy = (short)atoi(value);
        else if (!strcmp(keyword, "EnergyROI"))
            sscanf(value, "%lf/%lf/%lf/%lf", &p->roi[0], &p->roi[1], &p->roi[2], &p->roi[3]);
    }
    fclose(fid);
    if (p->iwipe == 0) p->np = 0;   /* io_utils.c:293 */
    return 0;
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

/*
 * Model bodies: the numbers after the header, whitespace-separated, read in
 * parallel.  Each thread reads its share of the rest of the file into one
 * buffer; the buffer is split into one range per thread at whitespace, the
 * numbers of each range are counted, and each thread then converts its own
 * exactly: fast_number where it applies, else strtod, the conversion
 * fscanf's %lf uses (both correctly rounded, so the values are
 * bit-identical to the reference's).  *nvalid is the count of leading
 * numbers before the first token strtod does not take whole (or the end);
 * out receives the first min(nvalid, want) of them.
 */
static int is_space(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

/* Decimal number at p in the common form [sign] digits [. digits]
 * [e|E [sign] digits], converted exactly without strtod when its digits
 * make an integer M <= 2^53 and its power of ten 10^E has |E| <= 22: both
 * are then exact doubles and one IEEE multiply or divide gives the
 * correctly rounded value (Clinger's fast path), the same as strtod.
 * Returns the end of the number, or NULL when the text is not in that form
 * or outside that range (the caller then uses strtod). */
static const char *fast_number(const char *p, double *out)
{
    static const double p10[23] = {1e0, 1e1, 1e2, 1e3, 1e4, 1e5, 1e6, 1e7, 1e8, 1e9, 1e10, 1e11,
                                   1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21,
                                   1e22};
    int neg = 0, nd = 0, any = 0;
    long e10 = 0;
    uint64_t m = 0;
    if (*p == '-' || *p == '+') neg = *p++ == '-';
    for (; *p >= '0' && *p <= '9'; ++p) {
        any = 1;
        if (m == 0 && *p == '0') continue;           /* leading zeros */
        if (++nd > 19) return NULL;
        m = m * 10 + (uint64_t)(*p - '0');
    }
    if (*p == '.') {
        for (++p; *p >= '0' && *p <= '9'; ++p) {
            any = 1;
            if (m == 0 && *p == '0') { --e10; continue; }
            if (++nd > 19) return NULL;
            m = m * 10 + (uint64_t)(*p - '0');
            --e10;
        }
    }
    if (!any) return NULL;
    if (*p == 'e' || *p == 'E') {
        const char *q = p + 1;
        int eneg = 0;
        long x = 0;
        if (*q == '-' || *q == '+') eneg = *q++ == '-';
        if (!(*q >= '0' && *q <= '9')) return NULL;
        for (; *q >= '0' && *q <= '9'; ++q)
            if ((x = x * 10 + (*q - '0')) > 100000) return NULL;
        e10 += eneg ? -x : x;
        p = q;
    }
    if (m > ((uint64_t)1 << 53)) return NULL;
    double v = (double)m;
    if (m != 0) {
        if (e10 < -22 || e10 > 22) return NULL;
        v = e10 < 0 ? v / p10[-e10] : v * p10[e10];
    }
    *out = neg ? -v : v;
    return p;
}

static double *read_body(FILE *f, const char *fn, size_t want, size_t *nvalid)
{
    const long start = ftell(f);
    if (start < 0 || fseek(f, 0, SEEK_END) != 0) die("Error: cannot read file %s\n", fn);
    const long end = ftell(f);
    if (end < start || fseek(f, start, SEEK_SET) != 0) die("Error: cannot read file %s\n", fn);
    const size_t len = (size_t)(end - start);
    char *buf = malloc(len + 1);
    double *out = malloc((want ? want : 1) * sizeof(double));
    if (!buf || !out) die("Error: out of memory reading %s\n", fn);
    buf[len] = '\0';

    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    if ((size_t)nt > len / 65536 + 1) nt = (int)(len / 65536 + 1);
    /* each thread reads its share of the file */
    int rd_ok = 1;
    const int fd = fileno(f);
#pragma omp parallel for num_threads(nt) schedule(static, 1) reduction(&& : rd_ok)
    for (int t = 0; t < nt; ++t) {
        size_t a = len / nt * t, b = t + 1 == nt ? len : len / nt * (t + 1);
        while (a < b) {
            const ssize_t r = pread(fd, buf + a, b - a, (off_t)start + (off_t)a);
            if (r <= 0) { rd_ok = 0; break; }
            a += (size_t)r;
        }
    }
    if (!rd_ok) die("Error: cannot read file %s\n", fn);
    size_t *lo = malloc((nt + 1) * sizeof(size_t)), *cnt = calloc(nt + 1, sizeof(size_t));
    size_t *bad = malloc(nt * sizeof(size_t));
    if (!lo || !cnt || !bad) die("Error: out of memory reading %s\n", fn);
    /* range t starts at a whitespace character (or 0) and ends where t+1 starts */
    lo[0] = 0;
    lo[nt] = len;
    for (int t = 1; t < nt; ++t) {
        size_t q = len / nt * t;
        if (q < lo[t - 1]) q = lo[t - 1];
        while (q < len && !is_space(buf[q])) ++q;
        lo[t] = q;
    }
    /* numbers (tokens) whose first character lies in each range */
#pragma omp parallel for num_threads(nt) schedule(static, 1)
    for (int t = 0; t < nt; ++t) {
        size_t c = 0;
        for (size_t q = lo[t]; q < lo[t + 1]; ++q)
            if (!is_space(buf[q]) && (q == 0 || is_space(buf[q - 1]))) ++c;
        cnt[t + 1] = c;
    }
    for (int t = 0; t < nt; ++t) cnt[t + 1] += cnt[t];   /* first token index of range t+1 */
#pragma omp parallel for num_threads(nt) schedule(static, 1)
    for (int t = 0; t < nt; ++t) {
        size_t k = cnt[t];
        bad[t] = (size_t)-1;
        size_t q = lo[t];
        while (q < lo[t + 1]) {
            if (is_space(buf[q])) { ++q; continue; }
            double v;
            const char *e = fast_number(buf + q, &v);
            if (!e || (*e != '\0' && !is_space(*e))) {
                char *e2;
                v = strtod(buf + q, &e2);
                e = e2;
            }
            if (e == buf + q || (*e != '\0' && !is_space(*e))) { bad[t] = k; break; }
            if (k < want) out[k] = v;
            ++k;
            q = (size_t)(e - buf);
        }
    }
    size_t n = cnt[nt];
    for (int t = 0; t < nt; ++t)
        if (bad[t] < n) n = bad[t];
    *nvalid = n;
    free(lo);
    free(cnt);
    free(bad);
    free(buf);
    return out;
}

/*
 * Model writers: the text read_model_e / read_model_a expect (a header of
 * dimensions, spacings, origin and model type, then one line per cell), with
 * every value printed with `digits` significant digits (17 makes the text
 * read back to the very same doubles).  Cells are formatted in parallel in
 * blocks and written in order.
 */
static int write_model_body(FILE *f, size_t ncell, int nval, const double *const *vals, int digits)
{
    /* 1..17 significant digits: a value then takes at most 24 characters
     * plus its separator, inside the 32 each one is given */
    if (digits < 1) digits = 1;
    if (digits > 17) digits = 17;
    const size_t block = (size_t)1 << 20;
    const size_t line_max = (size_t)nval * 32 + 2;
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    char *buf = malloc(block * line_max);
    size_t *len = malloc((size_t)nt * sizeof(size_t));
    if (!buf || !len) { free(buf); free(len); return -1; }
    int rc = 0;
    for (size_t c0 = 0; c0 < ncell && rc == 0; c0 += block) {
        const size_t nb = ncell - c0 < block ? ncell - c0 : block;
#pragma omp parallel for num_threads(nt) schedule(static, 1)
        for (int t = 0; t < nt; ++t) {
            const size_t a = nb * t / nt, b = nb * (t + 1) / nt;
            char *o = buf + a * line_max;
            size_t n = 0;
            for (size_t c = a; c < b; ++c)
                for (int v = 0; v < nval; ++v)
                    n += (size_t)snprintf(o + n, 32, v + 1 < nval ? "%.*g " : "%.*g\n", digits,
                                          vals[v][c0 + c]);
            len[t] = n;
        }
        for (int t = 0; t < nt && rc == 0; ++t)
            if (len[t] && fwrite(buf + nb * t / nt * line_max, 1, len[t], f) != len[t]) rc = -1;
    }
    free(buf);
    free(len);
    return rc;
}

int sm_write_model_a(const char *fn, size_t nx, size_t ny, size_t nz, double dx, double dy,
                     double dz, double x0, double y0, double z0, const double *vp,
                     const double *rho, int digits)
{
    /* the format io_utils.c:1029-1072 reads; vp, rho: nx*ny*nz, z fastest */
    FILE *f = fopen(fn, "w");
    if (!f) return -1;
    const double *vals[2] = {vp, rho};
    int rc = fprintf(f, "%zu %zu %zu\n%.17g %.17g %.17g\n%.17g %.17g %.17g\nisotropic_acoustic_3d\n",
                     nx, ny, nz, dx, dy, dz, x0, y0, z0) < 0 ? -1 : 0;
    if (rc == 0) rc = write_model_body(f, nx * ny * nz, 2, vals, digits);
    if (fclose(f) != 0) rc = -1;
    return rc;
}

int sm_write_model_e(const char *fn, size_t nx, size_t nz, double dx, double dz, double x0,
                     double z0, const double *vp, const double *vs, const double *rho, int digits)
{
    /* the format io_utils.c:892-947 reads; vp, vs, rho: nx*nz, z fastest */
    FILE *f = fopen(fn, "w");
    if (!f) return -1;
    const double *vals[3] = {vp, vs, rho};
    int rc = fprintf(f, "%zu %zu\n%.17g %.17g\n%.17g %.17g\nisotropic_elastic\n", nx, nz, dx, dz,
                     x0, z0) < 0 ? -1 : 0;
    if (rc == 0) rc = write_model_body(f, nx * nz, 3, vals, digits);
    if (fclose(f) != 0) rc = -1;
    return rc;
}

void sm_read_model_e(const char *fn, const sm_grid *g, double *vp, double *vs, double *rho)
{
    /* io_utils.c:892-947.  The reference reads nx columns of nz "Vp Vs rho"
     * triples with fscanf; when nothing can be read where the second column
     * starts, the file holds one column, which is copied to every x.  Any
     * other shortfall is an error. */
    FILE *f = open_or_die(fn);
    char type[64];
    if (fscanf(f, "%*d %*d") < 0 || fscanf(f, "%*f %*f") < 0 || fscanf(f, "%*f %*f") < 0 ||
        fscanf(f, "%63s", type) != 1 || strcmp(type, "isotropic_elastic") != 0)
        die("Error, model in %s not isotropic_elastic\n", fn);
    const size_t ncell = g->nx * g->nz;
    size_t nv = 0;
    double *v = read_body(f, fn, 3 * ncell, &nv);
    fclose(f);
    int flag1d = 0;
    if (nv < 3 * ncell) {
        if (g->nx > 1 && nv == 3 * g->nz) flag1d = 1;
        else die("Error: something wrong with file %s\n", fn);
    }
    const size_t ncopy = flag1d ? g->nz : ncell;
    for (size_t c = 0; c < ncopy; ++c) {
        vp[c] = v[3 * c];
        vs[c] = v[3 * c + 1];
        rho[c] = v[3 * c + 2];
    }
    free(v);
    if (flag1d)
        for (size_t i = 1; i < g->nx; ++i)
            for (size_t j = 0; j < g->nz; ++j) {
                vp[i * g->nz + j] = vp[j];
                vs[i * g->nz + j] = vs[j];
                rho[i * g->nz + j] = rho[j];
            }
}

void sm_read_grid_ve_sh(const char *fn, sm_grid *g)
{
    /* io_utils.c:330-354: the keyword compared in lower case */
    FILE *f = open_or_die(fn);
    memset(g, 0, sizeof(*g));
    char type[64];
    if (fscanf(f, "%zu %zu", &g->nx, &g->nz) != 2 || fscanf(f, "%lf %lf", &g->dx, &g->dz) != 2 ||
        fscanf(f, "%lf %lf", &g->x0, &g->z0) != 2 || fscanf(f, "%63s", type) != 1)
        die("Error: something wrong with file %s\n", fn);
    for (char *c = type; *c; ++c) *c = (char)tolower((unsigned char)*c);
    if (strcmp(type, "vti_sh_viscoelastic") != 0)
        die("Error, model file %s not for a viscoelastic VTI medium\n", fn);
    g->ny = 1;
    g->dy = g->dx;
    fclose(f);
}

void sm_read_model_ve_sh(const char *fn, const sm_grid *g, size_t np, double *mu, double *rho,
                         double *gamma, double *Q, double *fr)
{
    /* io_utils.c:1316-1429: "Vs0 rho gamma Q f" per node, x outer and z
     * inner, mu = Vs0^2 rho 1e-9 (GPa); one z column (copied to every x)
     * when nothing can be read where column 1 starts; values copied into the
     * absorbing layers from the nearest model node.  Any other shortfall is
     * an error. */
    FILE *f = open_or_die(fn);
    char type[64];
    if (fscanf(f, "%*d %*d") < 0 || fscanf(f, "%*f %*f") < 0 || fscanf(f, "%*f %*f") < 0 ||
        fscanf(f, "%63s", type) != 1 || strcmp(type, "vti_sh_viscoelastic") != 0)
        die("Error, model in %s not vti_sh_viscoelastic\n", fn);
    const size_t nx = g->nx, nz = g->nz, nz2 = nz + 2 * np;
    const size_t ncell = nx * nz;
    size_t nv = 0;
    double *v = read_body(f, fn, 5 * ncell, &nv);
    fclose(f);
    int flag1d = 0;
    if (nv < 5 * ncell) {
        if (nx > 1 && nv == 5 * nz) flag1d = 1;
        else die("Error: something wrong with file %s\n", fn);
    }
    for (size_t i = 0; i < nx; ++i)
        for (size_t j = 0; j < nz; ++j) {
            const double *c = v + 5 * ((flag1d ? 0 : i) * nz + j);
            const size_t ind = (i + np) * nz2 + j + np;
            rho[ind] = c[1];
            mu[ind] = c[0] * c[0] * rho[ind] * 1.e-9;
            gamma[ind] = c[2];
            Q[ind] = c[3];
            fr[ind] = c[4];
        }
    free(v);
    double *a[5] = {mu, rho, gamma, Q, fr};
    for (size_t i = 0; i < nx; ++i)
        for (size_t j = 0; j < np; ++j)
            for (int q = 0; q < 5; ++q) {
                a[q][(i + np) * nz2 + j] = a[q][(i + np) * nz2 + np];
                a[q][(i + np) * nz2 + np + nz + j] = a[q][(i + np) * nz2 + np + nz - 1];
            }
    for (size_t i = 0; i < np; ++i)
        for (size_t j = 0; j < nz2; ++j)
            for (int q = 0; q < 5; ++q) {
                a[q][i * nz2 + j] = a[q][np * nz2 + j];
                a[q][(i + nx + np) * nz2 + j] = a[q][(nx + np - 1) * nz2 + j];
            }
}

void sm_read_grid_ve(const char *fn, sm_grid *g, int *L)
{
    /* io_utils.c:356-382: the keyword compared in lower case, then L */
    FILE *f = open_or_die(fn);
    memset(g, 0, sizeof(*g));
    char type[64];
    if (fscanf(f, "%zu %zu", &g->nx, &g->nz) != 2 || fscanf(f, "%lf %lf", &g->dx, &g->dz) != 2 ||
        fscanf(f, "%lf %lf", &g->x0, &g->z0) != 2 || fscanf(f, "%63s", type) != 1)
        die("Error: something wrong with file %s\n", fn);
    for (char *c = type; *c; ++c) *c = (char)tolower((unsigned char)*c);
    if (strcmp(type, "vti_viscoelastic") != 0)
        die("Error, model file %s not for a viscoelastic VTI medium\n", fn);
    if (fscanf(f, "%d", L) != 1 || *L < 1)
        die("Error: something wrong with file %s\n", fn);
    g->ny = 1;
    g->dy = g->dx;
    fclose(f);
}

void sm_read_model_ve(const char *fn, const sm_grid *g, size_t np, int L, double *K, double *mu,
                      double *rho, double *eps, double *delta, double *Q1, double *f1, double *Q2,
                      double *f2)
{
    /* io_utils.c:1155-1314: per node "Vp0 Vs0 rho epsilon delta", L pairs
     * "Q1 f1", then "Q2 f2", x outer and z inner; mu = Vs0^2 rho 1e-9 and
     * K = 1e-9 (Vp0^2 rho - 4/3 Vs0^2 rho) in GPa; one z column (copied to
     * every x) when nothing can be read where column 1 starts; values
     * copied into the absorbing layers from the nearest model node.  Any
     * other shortfall is an error. */
    FILE *f = open_or_die(fn);
    char type[64];
    if (fscanf(f, "%*d %*d") < 0 || fscanf(f, "%*f %*f") < 0 || fscanf(f, "%*f %*f") < 0 ||
        fscanf(f, "%63s", type) != 1 || strcmp(type, "vti_viscoelastic") != 0)
        die("Error, model in %s not vti_viscoelastic\n", fn);
    if (fscanf(f, "%*d") < 0) die("Error: something wrong with file %s\n", fn);
    const size_t nx = g->nx, nz = g->nz, nz2 = nz + 2 * np, N = (nx + 2 * np) * nz2;
    const size_t per = 7 + 2 * (size_t)L, ncell = nx * nz;
    size_t nv = 0;
    double *v = read_body(f, fn, per * ncell, &nv);
    fclose(f);
    int flag1d = 0;
    if (nv < per * ncell) {
        if (nx > 1 && nv == per * nz) flag1d = 1;
        else die("Error: something wrong with file %s\n", fn);
    }
    for (size_t i = 0; i < nx; ++i)
        for (size_t j = 0; j < nz; ++j) {
            const double *c = v + per * ((flag1d ? 0 : i) * nz + j);
            const size_t ind = (i + np) * nz2 + j + np;
            const double vp0 = c[0], vs0 = c[1];
            rho[ind] = c[2];
            eps[ind] = c[3];
            delta[ind] = c[4];
            const double mu_pa = vs0 * vs0 * rho[ind];
            K[ind] = 1.e-9 * (vp0 * vp0 * rho[ind] - 4. / 3. * mu_pa);
            mu[ind] = mu_pa * 1.e-9;
            for (int m = 0; m < L; ++m) {
                Q1[(size_t)m * N + ind] = c[5 + 2 * m];
                f1[(size_t)m * N + ind] = c[6 + 2 * m];
            }
            Q2[ind] = c[5 + 2 * L];
            f2[ind] = c[6 + 2 * L];
        }
    free(v);
    const int na = 7 + 2 * L;
    double **a = malloc((size_t)na * sizeof(double *));
    a[0] = K; a[1] = mu; a[2] = rho; a[3] = eps; a[4] = delta; a[5] = Q2; a[6] = f2;
    for (int m = 0; m < L; ++m) { a[7 + 2 * m] = Q1 + (size_t)m * N; a[8 + 2 * m] = f1 + (size_t)m * N; }
    for (size_t i = 0; i < nx; ++i)
        for (size_t j = 0; j < np; ++j)
            for (int q = 0; q < na; ++q) {
                a[q][(i + np) * nz2 + j] = a[q][(i + np) * nz2 + np];
                a[q][(i + np) * nz2 + np + nz + j] = a[q][(i + np) * nz2 + np + nz - 1];
            }
    for (size_t i = 0; i < np; ++i)
        for (size_t j = 0; j < nz2; ++j)
            for (int q = 0; q < na; ++q) {
                a[q][i * nz2 + j] = a[q][np * nz2 + j];
                a[q][(i + nx + np) * nz2 + j] = a[q][(nx + np - 1) * nz2 + j];
            }
    free(a);
}

void sm_read_grid_pve(const char *fn, size_t np, sm_grid *g, int *with_pml)
{
    /* io_utils.c:296-328: the keyword compared in lower case; with the
     * layers in the file, nx and nz count them */
    FILE *f = open_or_die(fn);
    memset(g, 0, sizeof(*g));
    char type[64];
    if (fscanf(f, "%zu %zu", &g->nx, &g->nz) != 2 || fscanf(f, "%lf %lf", &g->dx, &g->dz) != 2 ||
        fscanf(f, "%lf %lf", &g->x0, &g->z0) != 2 || fscanf(f, "%63s", type) != 1)
        die("Error: something wrong with file %s\n", fn);
    fclose(f);
    for (char *c = type; *c; ++c) *c = (char)tolower((unsigned char)*c);
    *with_pml = strcmp(type, "isotropic_pml") == 0;
    if (!*with_pml && strcmp(type, "isotropic") != 0) die("Error, model in %s not isotropic\n", fn);
    if (*with_pml) {
        if (g->nx <= 2 * np || g->nz <= 2 * np)
            die("Error: the padded model in %s is not larger than its absorbing layers\n", fn);
        g->nx -= 2 * np;
        g->nz -= 2 * np;
    }
    g->ny = 1;
    g->dy = g->dx;
}

/* ncol values per node into cols[0..ncol), padded (nx + 2np) x (nz + 2np):
 * from the whole padded grid in the file (with_pml), or from nx x nz nodes
 * (or one z column: laterally uniform) copied into the layers from the
 * nearest model node */
static void read_pve_columns(const char *fn, const sm_grid *g, size_t np, int with_pml, int ncol,
                             double *const cols[])
{
    FILE *f = open_or_die(fn);
    char type[64];
    if (fscanf(f, "%*d %*d") < 0 || fscanf(f, "%*f %*f") < 0 || fscanf(f, "%*f %*f") < 0 ||
        fscanf(f, "%63s", type) != 1)
        die("Error: something wrong with file %s\n", fn);
    const size_t nx = g->nx, nz = g->nz, nx2 = nx + 2 * np, nz2 = nz + 2 * np, nc = (size_t)ncol;
    if (with_pml) {
        const size_t N = nx2 * nz2;
        size_t nv = 0;
        double *v = read_body(f, fn, nc * N, &nv);
        fclose(f);
        if (nv < nc * N) die("Error: something wrong with file %s\n", fn);
        for (size_t n = 0; n < N; ++n)
            for (size_t q = 0; q < nc; ++q) cols[q][n] = v[nc * n + q];
        free(v);
        return;
    }
    const size_t ncell = nx * nz;
    size_t nv = 0;
    double *v = read_body(f, fn, nc * ncell, &nv);
    fclose(f);
    int flag1d = 0;
    if (nv < nc * ncell) {
        if (nx > 1 && nv == nc * nz) flag1d = 1;
        else die("Error: something wrong with file %s\n", fn);
    }
    for (size_t i = 0; i < nx; ++i)
        for (size_t j = 0; j < nz; ++j) {
            const double *c = v + nc * ((flag1d ? 0 : i) * nz + j);
            for (size_t q = 0; q < nc; ++q) cols[q][(i + np) * nz2 + j + np] = c[q];
        }
    free(v);
    for (size_t i = 0; i < nx; ++i)
        for (size_t j = 0; j < np; ++j)
            for (size_t q = 0; q < nc; ++q) {
                double *a = cols[q];
                a[(i + np) * nz2 + j] = a[(i + np) * nz2 + np];
                a[(i + np) * nz2 + np + nz + j] = a[(i + np) * nz2 + np + nz - 1];
            }
    for (size_t i = 0; i < np; ++i)
        for (size_t j = 0; j < nz2; ++j)
            for (size_t q = 0; q < nc; ++q) {
                double *a = cols[q];
                a[i * nz2 + j] = a[np * nz2 + j];
                a[(i + nx + np) * nz2 + j] = a[(nx + np - 1) * nz2 + j];
            }
}

void sm_read_model_pve(const char *fn, const sm_grid *g, size_t np, int with_pml, double *const vals[12])
{
    /* io_utils.c:403-625.  "isotropic": a short file holding exactly one z
     * column is laterally uniform, any other shortfall an error; the layers
     * take the nearest model node's values (Qdamping is never set, so Q is
     * copied too).  "isotropic_pml": every padded node from the file (the
     * reference leaves a short last node uninitialised; here it is an
     * error). */
    read_pve_columns(fn, g, np, with_pml, 12, vals);
}

void sm_read_grid_pve_vti(const char *fn, sm_grid *g, int *with_pml)
{
    /* io_utils.c:296-328 and 627-647: read_grid_params subtracts the layers
     * only for "isotropic_pml", so nx and nz of an "anisotropic__vti_pml"
     * file (two underscores, as read_modelVTI compares it) count interior
     * nodes while its body holds the padded grid */
    FILE *f = open_or_die(fn);
    memset(g, 0, sizeof(*g));
    char type[64];
    if (fscanf(f, "%zu %zu", &g->nx, &g->nz) != 2 || fscanf(f, "%lf %lf", &g->dx, &g->dz) != 2 ||
        fscanf(f, "%lf %lf", &g->x0, &g->z0) != 2 || fscanf(f, "%63s", type) != 1)
        die("Error: something wrong with file %s\n", fn);
    fclose(f);
    for (char *c = type; *c; ++c) *c = (char)tolower((unsigned char)*c);
    *with_pml = strcmp(type, "anisotropic__vti_pml") == 0;
    if (!*with_pml && strcmp(type, "anisotropic_vti") != 0) die("Error, model in %s not VTI anisotropic\n", fn);
    g->ny = 1;
    g->dy = g->dx;
}

void sm_read_model_pve_vti(const char *fn, const sm_grid *g, size_t np, int with_pml, double *const vals[16])
{
    /* io_utils.c:627-888: per node "K_m K_s K_f phi mu rho_s rho_f T3 eta
     * kappa3 Q f0 epsilon delta kappa3/kappa1 T3/T1", into vals in the order
     * K_m K_s K_f phi mu rho_s rho_f T1 T3 eta kappa1 kappa3 Q f0 epsilon
     * delta; the same column and layer rules as sm_read_model_pve */
    double *const cols[16] = {vals[0], vals[1], vals[2], vals[3], vals[4], vals[5], vals[6], vals[8],
                              vals[9], vals[11], vals[12], vals[13], vals[14], vals[15], vals[10], vals[7]};
    read_pve_columns(fn, g, np, with_pml, 16, cols);
    const size_t N = (g->nx + 2 * np) * (g->nz + 2 * np);
    for (size_t n = 0; n < N; ++n) {
        vals[10][n] = vals[11][n] / vals[10][n];   /* kappa1 = kappa3 / (kappa3/kappa1) */
        vals[7][n] = vals[8][n] / vals[7][n];      /* T1 = T3 / (T3/T1) */
    }
}

void sm_pve_dispersion(size_t n, double *const v[12], double f, double *const vel[3], double *const att[3])
{
    /* the node material in the units of the time loop (pve_iso_pml.c:795-
     * 829: MPa, Mg/m^3, eta scaled so that eta/kappa is in Mg/m^3/s), then
     * the complex density rho_bar = T rho_f / phi - i (eta/kappa) / omega
     * and fluid modulus M (1 + i omega tau_e) / ((1 + varphi)(1 + i omega
     * tau_s)), and the roots of the dispersion relation */
    const double pi = 4.0 * atan(1.0);
    const double omega = 2. * pi * f;
    for (size_t c = 0; c < n; ++c) {
        const double K_m = v[0][c] * 1000., K_s = v[1][c] * 1000., K_f = v[2][c] * 1000., phi = v[3][c];
        const double mu = v[4][c] * 1000., rho_s = v[5][c] * 1.e-6, rho_f = v[6][c] * 1.e-6, T = v[7][c];
        const double eta = v[8][c] * 1.013249965828145e6, kappa = v[9][c], Q = v[10][c];
        const double tau0 = 1. / (2. * pi * v[11][c]);
        const double E = K_m + 4. / 3. * mu;
        const double D = K_s * (1. + phi * (K_s / K_f - 1.));
        const double M = K_s * K_s / (D - K_m);
        const double alpha = 1. - K_m / K_s;
        double tau_e = tau0, tau_s = tau0, varphi = 0.0;
        if (Q < 1.e4) {
            tau_e = tau0 / Q * (sqrt(Q * Q + 1.) + 1.);
            tau_s = tau0 / Q * (sqrt(Q * Q + 1.) - 1.);
            varphi = tau_e / tau_s - 1.;
        }
        const double rho = (1. - phi) * rho_s + phi * rho_f;
        const double complex rho_bar = T / phi * rho_f - I / omega * (eta / kappa);
        const double complex rho_c = rho - rho_f * rho_f / rho_bar;
        const double complex M_c = M / (1. + varphi) * (1. + I * omega * tau_e) / (1. + I * omega * tau_s);
        const double complex A = M_c * (rho - 2. * alpha * rho_f) + rho_bar * (E + alpha * alpha * M_c);
        const double complex tmp = csqrt(A * A - 4. * M_c * E * rho_c * rho_bar);
        const double complex V[3] = {csqrt((A - tmp) / (2. * rho_c * rho_bar)),
                                     csqrt((A + tmp) / (2. * rho_c * rho_bar)), csqrt(mu / rho_c)};
        for (int k = 0; k < 3; ++k) {
            vel[k][c] = 1. / creal(1. / V[k]);
            att[k][c] = 17.372 * pi * cimag(V[k]) / creal(V[k]);
        }
    }
}

/* phase velocities of the fast (vf) and slow (vs) P waves of a Biot medium
 * along one axis: det(C - V^2 R) = 0 with the undrained stiffness
 * C = [[c11, c12], [c12, c22]] and the density matrix R = [[rho, rho_f],
 * [rho_f, Y]], Y = m - i (eta/kappa)/omega */
static void biot_axis(double c11, double c12, double c22, double rho, double rho_f, double complex Y,
                      double complex *vf, double complex *vs)
{
    const double complex a = rho * Y - rho_f * rho_f;
    const double complex b = -(c11 * Y + c22 * rho - 2. * c12 * rho_f);
    const double c = c11 * c22 - c12 * c12;
    const double complex d = csqrt(b * b - 4. * a * c);
    const double complex v1 = csqrt((-b + d) / (2. * a)), v2 = csqrt((-b - d) / (2. * a));
    if (creal(v1) >= creal(v2)) { *vf = v1; *vs = v2; }
    else { *vf = v2; *vs = v1; }
}

void sm_pve_dispersion_vti(size_t nx2, size_t nz2, double *const v[16], double f, double *const vel[6],
                           double *const att[6])
{
    /* the VTI medium's node material (pve_vti_pml.c:787-829), the
     * permeability terms and c55 at their staggered positions as the
     * reference has them when it runs the check (837-929), the undrained
     * stiffness c + alpha alpha M, and plane waves along x and z: qP from
     * biot_axis with the axis' stiffness, alpha, tortuosity and
     * permeability; qS polarised across the axis, V^2 = c55 / (rho -
     * rho_f^2 / Y) with the Y of the polarisation axis */
    const double pi = 4.0 * atan(1.0), omega = 2. * pi * f;
    const size_t N = nx2 * nz2;
    for (size_t c = 0; c < N; ++c) {
        const size_t i = c / nz2, j = c % nz2;
        const double K_m = v[0][c] * 1000., K_s = v[1][c] * 1000., K_f = v[2][c] * 1000., phi = v[3][c];
        const double mu = v[4][c] * 1000., rho_s = v[5][c] * 1.e-6, rho_f = v[6][c] * 1.e-6;
        const double D = K_s * (1. + phi * (K_s / K_f - 1.));
        const double c33 = K_m + 4. / 3. * mu, c11 = 2. * v[14][c] * c33 + c33, c12 = c11 - 2. * mu;
        const double c13 = sqrt(2. * v[15][c] * c33 * (c33 - mu) + (c33 - mu) * (c33 - mu)) - mu;
        const double M = K_s * K_s / (D - (2. * c11 + c33 + 2. * c12 + 4. * c13) / 9.);
        const double a1 = 1. - 1. / 3. * (c11 + c12 + c13) / K_s, a3 = 1. - 1. / 3. * (2. * c13 + c33) / K_s;
        const double rho = (1. - phi) * rho_s + phi * rho_f;
        const double m1 = v[7][c] * rho_f / phi, m3 = v[8][c] * rho_f / phi;
        /* eta / kappa averaged to (i+1/2, j) and (i, j+1/2), the last row or
         * column copied from its neighbour */
        const size_t ia = i + 1 < nx2 ? i : i - 1, ja = j + 1 < nz2 ? j : j - 1;
        const size_t a = ia * nz2 + j, b = a + nz2, a3i = i * nz2 + ja, b3 = a3i + 1;
        const double E = 1.013249965828145e6;
        const double nk1 = 0.5 * (v[9][a] * E / v[10][a] + v[9][b] * E / v[10][b]);
        const double nk3 = 0.5 * (v[9][a3i] * E / v[11][a3i] + v[9][b3] * E / v[11][b3]);
        /* harmonic mean of mu at (i+1/2, j+1/2), last row and column copied */
        const size_t ic = i + 1 < nx2 ? i : i - 1, jc = j + 1 < nz2 ? j : j - 1;
        const size_t q00 = ic * nz2 + jc;
        const double c55 = 4.0 / (1. / (v[4][q00] * 1000.) + 1. / (v[4][q00 + nz2] * 1000.) +
                                  1. / (v[4][q00 + 1] * 1000.) + 1. / (v[4][q00 + nz2 + 1] * 1000.));
        const double complex Y1 = m1 - I * nk1 / omega, Y3 = m3 - I * nk3 / omega;
        double complex V[6];
        biot_axis(c11 + a1 * a1 * M, a1 * M, M, rho, rho_f, Y1, &V[0], &V[1]);
        V[2] = csqrt(c55 / (rho - rho_f * rho_f / Y3));
        biot_axis(c33 + a3 * a3 * M, a3 * M, M, rho, rho_f, Y3, &V[3], &V[4]);
        V[5] = csqrt(c55 / (rho - rho_f * rho_f / Y1));
        for (int k = 0; k < 6; ++k) {
            vel[k][c] = 1. / creal(1. / V[k]);
            att[k][c] = -omega * cimag(1. / V[k]);
        }
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
    /* "Vp rho" pairs, x and y outer, z inner.  As io_utils.c:1049-1072
     * reads them, the model is one z column (copied everywhere) when nothing
     * can be read where column (i, 1) of any x plane i starts, that is when
     * the numbers end right after column (i, 0); any other shortfall is an
     * error. */
    const size_t ncell = g->nx * g->ny * g->nz;
    size_t nv = 0;
    double *v = read_body(f, fn, 2 * ncell, &nv);
    fclose(f);
    int flag1d = 0;
    if (nv < 2 * ncell) {
        if (g->ny > 1 && nv % (2 * g->nz * g->ny) == 2 * g->nz) flag1d = 1;
        else die("Error: something wrong with file %s\n", fn);
    }
    const size_t ncol = flag1d ? 1 : g->nx * g->ny;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < ncol; ++c) {
        const size_t i = c / g->ny, j = c % g->ny;
        for (size_t k = 0; k < g->nz; ++k) {
            const size_t q = c * g->nz + k;
            vp[AT(i + np, j + np, k + np)] = v[2 * q];
            rho[AT(i + np, j + np, k + np)] = v[2 * q + 1];
        }
    }
    free(v);
    if (flag1d)
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < g->nx; ++i)
            for (size_t j = i == 0 ? 1 : 0; j < g->ny; ++j)
                for (size_t k = 0; k < g->nz; ++k) {
                    vp[AT(i + np, j + np, k + np)] = vp[AT(np, np, k + np)];
                    rho[AT(i + np, j + np, k + np)] = rho[AT(np, np, k + np)];
                }
    /* edge values into the absorbing layers: z, then y, then x (each pass
     * reads only what the earlier passes wrote; its outer iterations write
     * disjoint cells) */
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < g->nx; ++i)
        for (size_t j = 0; j < g->ny; ++j)
            for (size_t k = 0; k < np; ++k) {
                vp[AT(i + np, j + np, k)] = vp[AT(i + np, j + np, np)];
                rho[AT(i + np, j + np, k)] = rho[AT(i + np, j + np, np)];
                vp[AT(i + np, j + np, np + g->nz + k)] = vp[AT(i + np, j + np, np + g->nz - 1)];
                rho[AT(i + np, j + np, np + g->nz + k)] = rho[AT(i + np, j + np, np + g->nz - 1)];
            }
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < g->nx; ++i)
        for (size_t k = 0; k < nz2; ++k)
            for (size_t j = 0; j < np; ++j) {
                vp[AT(i + np, j, k)] = vp[AT(i + np, np, k)];
                rho[AT(i + np, j, k)] = rho[AT(i + np, np, k)];
                vp[AT(i + np, np + g->ny + j, k)] = vp[AT(i + np, np + g->ny - 1, k)];
                rho[AT(i + np, np + g->ny + j, k)] = rho[AT(i + np, np + g->ny - 1, k)];
            }
#pragma omp parallel for schedule(static)
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

/* padded grid index round((x - x0) / d) + np as the reference computes it
 * (io_utils.c:1488-1492 with np = 0, 1582-1585); an index outside the
 * padded axis of n + 2 np nodes, where the reference would write out of
 * bounds, is an error.  3D positions inside the CPML layers are valid. */
static size_t node_padded(double x, double x0, double d, size_t n, size_t np, const char *what)
{
    const double r = round((x - x0) / d) + (double)np;
    if (!(r >= 0.0) || r >= (double)(n + 2 * np)) {
        fprintf(stderr, "Error: %s coordinate %g outside the grid\n", what, x);
        exit(1);
    }
    return (size_t)r;
}

static size_t node(double x, double x0, double d, size_t n, const char *what)
{
    return node_padded(x, x0, d, n, 0, what);
}

void sm_read_source2d(const char *fn, const sm_grid *g, sm_sources *src)
{
    sm_read_source2d_np(fn, g, 0, src);   /* e_iso: np = 0 */
}

void sm_read_source2d_np(const char *fn, const sm_grid *g, size_t np, sm_sources *src)
{
    /* io_utils.c:1432-1538: indices round((x - x0)/dx) + np on the padded
     * grid of nz + 2 np nodes along z */
    const size_t nz2 = g->nz + 2 * np;
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
        size_t i = node_padded(s->x, g->x0, g->dx, g->nx, np, "source x");
        size_t j = node_padded(s->z, g->z0, g->dz, g->nz, np, "source z");
        s->idx = i * nz2 + j;
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
                size_t ii = node_padded(s->x + di[q] * g->dx, g->x0, g->dx, g->nx, np, "template x");
                size_t jj = node_padded(s->z + dj[q] * g->dz, g->z0, g->dz, g->nz, np, "template z");
                e->idx = ii * nz2 + jj;
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
        size_t i = node_padded(s->x, g->x0, g->dx, g->nx, np, "source x");
        size_t j = node_padded(s->y, g->y0, g->dy, g->ny, np, "source y");
        size_t k = node_padded(s->z, g->z0, g->dz, g->nz, np, "source z");
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
            /* lround(...) + np, io_utils.c:1661-1666, 1722-1725: a 3D
             * record may sit in the CPML layers; outside the padded grid
             * the reference would read out of bounds */
            const double fi = (double)lround((r->x - g->x0) / g->dx) + (double)np;
            const double fk = (double)lround((r->z - g->z0) / g->dz) + (double)np;
            const double fj = three_d ? (double)lround((r->y - g->y0) / g->dy) + (double)np : 0.0;
            const double nx2 = (double)(g->nx + 2 * np), nz2 = (double)(g->nz + 2 * np),
                         ny2 = three_d ? (double)(g->ny + 2 * np) : 1.0;
            if (fi < 0 || fi >= nx2 || fk < 0 || fk >= nz2 || fj < 0 || fj >= ny2) {
                fprintf(stderr, "Error: record %zu is outside the grid\n", n + 1);
                exit(1);
            }
            const size_t i = (size_t)fi, j = (size_t)fj, k = (size_t)fk;
            r->idx = three_d ? (i * ny2 + j) * nz2 + k : i * nz2 + k;
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

void sm_read_output2d_np(const char *fn, const sm_grid *g, size_t np, sm_records *out)
{
    read_output(fn, g, np, out, 0);   /* io_utils.c:1592-1682 on the padded grid */
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

/* field: NULL for a snapshot (ssh/<base>/<base>_<nr+1>_<t>.nc, values in
 * mm/s), else write_field_nc's <base>_<field>.nc with the values unscaled
 * and the given units */
static int grid_nc(const double *data, double t, size_t nx, size_t nz, size_t np,
                   double x0, double z0, double dx, double dz, const char *basename,
                   size_t nr, int plstr, int cyl, const char *field, const char *units)
{
    /* write_snapshot_nc / write_snapshot3d_nc (io_utils.c:1822-2180): dims
     * x, y; double coordinate variables x(x), y(y); float z(y, x) in mm/s.
     * Cylindrical grids (cyl) have no layer at the axis: nx + np rows, the
     * x coordinates start at x0 (io_utils.c:1877-1881, 1946-1949); without
     * strips the reference still skips np rows (io_utils.c:1985-1991), and
     * so does this writer. */
    char fn[3 * SM_NAME_LEN + 64];
    const double scale = field ? 1.0 : 1.e3;
    if (field) {
        snprintf(fn, sizeof(fn), "%s_%s.nc", basename, field);
    } else {
        make_dir("ssh");
        snprintf(fn, sizeof(fn), "ssh/%s", basename);
        make_dir(fn);
        snprintf(fn, sizeof(fn), "ssh/%s/%s_%zu_%lf.nc", basename, basename, nr + 1, t);
    }
    const size_t nx2 = cyl ? nx + np : nx + 2 * np, nz2 = nz + 2 * np;
    const size_t mx = plstr ? nx2 : nx, mz = plstr ? nz2 : nz;
    const size_t ox = plstr ? 0 : np, oz = plstr ? 0 : np;
    double xr[2], yr[2], zr[2];
    if (plstr) {
        xr[0] = cyl ? x0 : x0 - np * dx; xr[1] = x0 + (nx + np - 1) * dx;
        yr[0] = z0 - np * dz; yr[1] = z0 + (nz + np - 1) * dz;
    } else {
        xr[0] = x0; xr[1] = x0 + (nx - 1) * dx;
        yr[0] = z0; yr[1] = z0 + (nz - 1) * dz;
    }
    zr[0] = zr[1] = scale * data[ox * nz2 + oz];
    for (size_t i = 0; i < mx; ++i)
        for (size_t j = 0; j < mz; ++j) {
            const double v = scale * data[(i + ox) * nz2 + j + oz];
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
            /* write_snapshot_nc counts the NUL of "mm/s", write_field_nc
             * gives strlen(units) */
            if (field) att_text(&h, "units", units, strlen(units));
            else att_text(&h, "units", "mm/s", 5);
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
        put_f64(&h, plstr && !cyl ? x0 + ((1.0 * i) - np) * dx : x0 + i * dx);
    for (size_t j = 0; j < mz; ++j)
        put_f64(&h, plstr ? z0 + ((1.0 * j) - np) * dz : z0 + j * dz);
    for (size_t j = 0; j < mz; ++j)
        for (size_t i = 0; i < mx; ++i)
            put_f32(&h, field ? (float)data[(i + ox) * nz2 + j + oz]
                              : (float)data[(i + ox) * nz2 + j + oz] * 1.e3f);
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

int sm_write_snapshot_nc(const double *data, double t, size_t nx, size_t nz, size_t np,
                         double x0, double z0, double dx, double dz, const char *basename,
                         size_t nr, int plstr)
{
    return grid_nc(data, t, nx, nz, np, x0, z0, dx, dz, basename, nr, plstr, 0, NULL, NULL);
}

int sm_write_field_nc(const double *data, const char *field, const char *units, size_t nx, size_t nz,
                      size_t np, double x0, double z0, double dx, double dz, const char *basename,
                      int plstr)
{
    return grid_nc(data, 0.0, nx, nz, np, x0, z0, dx, dz, basename, 0, plstr, 0, field, units);
}

int sm_write_snapshot_cyl_nc(const double *data, double t, size_t nx, size_t nz, size_t np,
                             double x0, double z0, double dx, double dz, const char *basename,
                             size_t nr, int plstr)
{
    return grid_nc(data, t, nx, nz, np, x0, z0, dx, dz, basename, nr, plstr, 1, NULL, NULL);
}

/* ------------------------------------------------------------------ */
/* e_cyl_iso inputs                                                   */
/* ------------------------------------------------------------------ */
void sm_read_model_e_cyl(const char *fn, const sm_grid *g, size_t np, double *vp, double *vs,
                         double *rho)
{
    /* read_model_e_cyl (io_utils.c:949-1027): after the header, "Vp Vs rho"
     * with r fastest, into rows i, columns j + np of the nx+np by nz+2np
     * grid; a file holding one row of r values (the read of (0, 1) fails)
     * is copied to every z; values are then copied into the z layers and
     * the outer r layer.  A file ending elsewhere is an error here (the
     * reference goes on with unset values). */
    FILE *f = open_or_die(fn);
    const size_t nx = g->nx, nz = g->nz, nz2 = nz + 2 * np, nx2 = nx + np;
    char type[64];
    if (fscanf(f, "%*d %*d") < 0 || fscanf(f, "%*f %*f") < 0 || fscanf(f, "%*f %*f") < 0 ||
        fscanf(f, "%63s", type) != 1)
        die("Error: something wrong with file %s\n", fn);
    if (strcmp(type, "isotropic_elastic_cyl") != 0) die("Error, model in %s not isotropic_elastic_cyl\n", fn);
    int column = 0;
    for (size_t j = 0; j < nz && !column; ++j)
        for (size_t i = 0; i < nx; ++i) {
            const size_t c = i * nz2 + j + np;
            const int nread = fscanf(f, "%lf %lf %lf", &vp[c], &vs[c], &rho[c]);
            if (nread <= 0 && i == 0 && j == 1) { column = 1; break; }
            if (nread != 3) die("Error: something wrong with file %s\n", fn);
        }
    fclose(f);
    if (column)
        for (size_t i = 0; i < nx; ++i)
            for (size_t j = 1; j < nz; ++j) {
                vp[i * nz2 + j + np] = vp[i * nz2 + np];
                vs[i * nz2 + j + np] = vs[i * nz2 + np];
                rho[i * nz2 + j + np] = rho[i * nz2 + np];
            }
    for (size_t i = 0; i < nx; ++i) {
        for (size_t j = 0; j < np; ++j) {
            vp[i * nz2 + j] = vp[i * nz2 + np];
            vs[i * nz2 + j] = vs[i * nz2 + np];
            rho[i * nz2 + j] = rho[i * nz2 + np];
        }
        for (size_t j = nz + np; j < nz2; ++j) {
            vp[i * nz2 + j] = vp[i * nz2 + nz + np - 1];
            vs[i * nz2 + j] = vs[i * nz2 + nz + np - 1];
            rho[i * nz2 + j] = rho[i * nz2 + nz + np - 1];
        }
    }
    for (size_t i = nx; i < nx2; ++i)
        for (size_t j = 0; j < nz2; ++j) {
            vp[i * nz2 + j] = vp[(nx - 1) * nz2 + j];
            vs[i * nz2 + j] = vs[(nx - 1) * nz2 + j];
            rho[i * nz2 + j] = rho[(nx - 1) * nz2 + j];
        }
}

void sm_read_source_cyl(const char *fn, const sm_grid *g, size_t np, sm_sources *src)
{
    /* read_source (io_utils.c:1432-1538) on a cylindrical grid: index
     * round((r - x0)/dr) * nz2 + round((z - z0)/dz) + np, the row without
     * the np offset (1489-1490).  The 9-point template keeps the reference's
     * offsets, which add np to the rows of its eight extra points
     * (1515-1534); they take the main point's type. */
    FILE *f = open_or_die(fn);
    memset(src, 0, sizeof(*src));
    if (fscanf(f, "%zu", &src->nsrc) != 1) die("Error: something wrong with file %s\n", fn);
    if (fscanf(f, "%zu", &src->ntemplate) != 1) src->ntemplate = 1;
    if (src->ntemplate != 1 && src->ntemplate != 9)
        die("Error: source template must have 1 or 9 points (%s)\n", fn);
    src->s = calloc(src->nsrc * src->ntemplate, sizeof(sm_source));
    const size_t nx2 = g->nx + np, nz2 = g->nz + 2 * np;
    char t[64];
    for (size_t n = 0; n < src->nsrc * src->ntemplate; n += src->ntemplate) {
        sm_source *s = &src->s[n];
        if (fscanf(f, "%63s", t) != 1) die("Error: something wrong with file %s\n", fn);
        s->type = source_type(t, 0);
        if (s->type < 0) die("Error: type of source (%s) not implemented\n", t);
        if (fscanf(f, "%lf", &s->A) != 1 || fscanf(f, "%lf", &s->f) != 1 ||
            fscanf(f, "%lf %lf", &s->x, &s->z) != 2)
            die("Error: something wrong with file %s\n", fn);
        const size_t i = node_padded(s->x, g->x0, g->dx, nx2, 0, "source r");
        const size_t j = node_padded(s->z, g->z0, g->dz, g->nz, np, "source z");
        s->idx = i * nz2 + j;
        if (src->ntemplate == 9) {
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
                if (q == 0) continue;
                const size_t ii = node_padded(s->x + di[q] * g->dx, g->x0, g->dx, nx2 - np, np,
                                              "template r");
                const size_t jj = node_padded(s->z + dj[q] * g->dz, g->z0, g->dz, g->nz, np,
                                              "template z");
                if (ii >= nx2) {
                    fprintf(stderr, "Error: template r coordinate outside the grid\n");
                    exit(1);
                }
                e->idx = ii * nz2 + jj;
            }
        }
    }
    fclose(f);
}

void sm_read_output_cyl(const char *fn, const sm_grid *g, size_t np, sm_records *out)
{
    /* read_output (io_utils.c:1592-1682) on a cylindrical grid: row
     * round((r - x0)/dr), column lround((z - z0)/dz) + np (1661-1666) */
    FILE *f = open_or_die(fn);
    if (fscanf(f, "%zu", &out->nrec) != 1) die("Error: something wrong with file %s\n", fn);
    out->r = calloc(out->nrec ? out->nrec : 1, sizeof(sm_record));
    const size_t nx2 = g->nx + np, nz2 = g->nz + 2 * np;
    char t[80];
    for (size_t n = 0; n < out->nrec; ++n) {
        sm_record *r = &out->r[n];
        if (fscanf(f, "%79s", t) != 1) die("Error: something wrong with file %s\n", fn);
        r->comp = comp_code(t, 0);
        if (r->comp < 0) die("Error: particle velocity component (%s) does not exists\n", t);
        if (fscanf(f, "%79s", t) != 1) die("Error: something wrong with file %s\n", fn);
        if (!strcmp(t, "Trace") || !strcmp(t, "trace") || !strcmp(t, "TRACE")) {
            r->type = REC_TRACE;
            if (fscanf(f, "%lf %lf", &r->x, &r->z) != 2 || fscanf(f, "%lf", &r->t0) != 1 ||
                fscanf(f, "%lf", &r->dt) != 1)
                die("Error: something wrong with file %s\n", fn);
            r->dt *= 1.e-3;
            const double fi = round((r->x - g->x0) / g->dx);
            const double fk = (double)lround((r->z - g->z0) / g->dz) + (double)np;
            if (!(fi >= 0) || fi >= (double)nx2 || !(fk >= 0) || fk >= (double)nz2) {
                fprintf(stderr, "Error: record %zu is outside the grid\n", n + 1);
                exit(1);
            }
            r->idx = (size_t)fi * nz2 + (size_t)fk;
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
