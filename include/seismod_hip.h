/*
 * seismod_hip.h -- C ABI of the MI355X FDTD library (libseismod_hip.so).
 *
 * The reference (seismod) has no library API: its boundary is two programs,
 * e_iso (2D elastic, split PML) and a_iso_3d (3D acoustic, CPML), whose time
 * loops call plain C functions or run inline loops.  This ABI replaces
 * those loops and their FP64 set-up; the drop-in programs in
 * seismod_amd/host/ keep the reference's files and command line and call it.
 *
 * Conventions
 *   - Every function returns an int status (SM_OK = 0); the library never
 *     exits or aborts.  sm_last_error() gives a message for the calling
 *     thread's last failure.
 *   - Host arrays are FP64 in the reference's layouts: 2D index i*nz+j and
 *     3D index (i*ny2+j)*nz2+k, z fastest.  Device layout and precision are
 *     private to the library.
 *   - One sm_ctx owns one HIP stream on one device.  Solvers created on a
 *     context run on its stream; different contexts run concurrently.
 *     A context and its solvers must not be used from two threads at once.
 *   - Enumerations reuse the reference's numbering (structs.h:44-45).
 */
#ifndef SEISMOD_HIP_H
#define SEISMOD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SM_ABI_VERSION 1

/* status codes */
#define SM_OK 0
#define SM_ERR_ARG 1     /* bad argument or size */
#define SM_ERR_HIP 2     /* HIP runtime error */
#define SM_ERR_CFL 3     /* time step violates the stability limit */
#define SM_ERR_NOMEM 4   /* host or device allocation failed */
#define SM_ERR_STATE 5   /* call not valid in the solver's current state */
#define SM_ERR_NODEV 6   /* no usable GPU */

/* wavefield storage formats (arithmetic is FP32 except SM_STORE_F64) */
#define SM_STORE_F32 0
#define SM_STORE_F64 1   /* FP64 storage and arithmetic, same operation
                            order as the reference */
#define SM_STORE_BF16 2  /* bfloat16 with a per-field power-of-two scale */
#define SM_STORE_F16 3   /* IEEE half with a per-field power-of-two scale */

/* record components, numbering of enum component (structs.h:45) */
#define SM_VX 0
#define SM_VY 1
#define SM_VZ 2
#define SM_QX 3
#define SM_QZ 4
#define SM_TXX 5
#define SM_TZZ 6
#define SM_TXZ 7
#define SM_P 8
#define SM_TXY 9
#define SM_TYZ 10
#define SM_DIV 11
#define SM_CURL 12
#define SM_VR 13
#define SM_VT 14
#define SM_TRR 15
#define SM_TRT 16
#define SM_TRZ 17
#define SM_TTT 18
#define SM_TTZ 19

/* source types, numbering of enum sourceType (structs.h:44) */
#define SM_SX 0
#define SM_SY 1
#define SM_SZ 2
#define SM_SXY 3
#define SM_SXZ 4
#define SM_SYZ 5
#define SM_BULK 6
#define SM_SF 7
#define SM_BULK_S 8
#define SM_FR 9
#define SM_FT 10
#define SM_FZ 11
#define SM_KURKJIAN 12

typedef struct sm_ctx sm_ctx;
typedef struct sm_a3d sm_a3d;
typedef struct sm_e2d sm_e2d;
typedef struct sm_cyl sm_cyl;

/* ------------------------------------------------------------------ */
/* library and context                                                 */
/* ------------------------------------------------------------------ */
int sm_abi_version(void);
const char *sm_strerror(int status);
/* message of the calling thread's last failed call ("" if none) */
const char *sm_last_error(void);

/* One context = one device + one HIP stream (default priority). */
int sm_ctx_create(int device, sm_ctx **out);
int sm_ctx_destroy(sm_ctx *ctx);
/* wait for all work queued on the context's stream */
int sm_ctx_sync(sm_ctx *ctx);
/* free and total device memory (hipMemGetInfo) of the context's device */
int sm_ctx_mem_info(sm_ctx *ctx, size_t *free_bytes, size_t *total_bytes);
/* Calibration: a plain streaming kernel reading `reads` and writing
 * `writes` FP32 arrays of n values (known traffic 4n(reads+writes) bytes),
 * each lane moving `vec` (1, 2 or 4) consecutive values per array; returns
 * the mean time of `iters` launches.  (reads, writes) is one of (1,1)
 * (2,1) (3,1) (4,2) (6,4) (8,4) (10,4) (12,8) (1,2) (5,5).  Optional bits
 * in `vec`: SM_PROBE_NT (non-temporal accesses), SM_PROBE_TILES (one
 * 256-element tile per workgroup instead of a persistent grid of 16
 * workgroups per CU).  No counterpart in the reference: it scales HBM
 * counters and measures achievable bandwidth. */
#define SM_PROBE_NT 16
#define SM_PROBE_TILES 32
int sm_probe_stream(sm_ctx *ctx, size_t n, int reads, int writes, int vec, int iters,
                    double *ms_per_iter);

/* ------------------------------------------------------------------ */
/* host-side FP64 set-up shared by both programs                       */
/* ------------------------------------------------------------------ */
/* Ricker length ceil(2/(f*dt)) -- replaces src.c:50-51. */
size_t sm_ricker_length(double f, double dt);
/* compute_src_fct (src.c:43-65) for n sources, fct concatenated (source s
 * starts at sum(length[0..s-1])); keeps the reference's normalisation of
 * every pass through source 0.  length[] from sm_ricker_length. */
int sm_compute_src_fct(size_t n, const double *A, const double *f, double dt,
                       const size_t *length, double *fct);
/* time-step count of e_iso.c:575-576 and a_iso_3d.c:578-579 */
size_t sm_nsteps(double duration, double dt);
/* record predicate of e_iso.c:643-658 / a_iso_3d.c:700-718: non-zero when
 * a record with start t0 [s] and interval rdt [s] samples after step it */
int sm_record_due(size_t it, double dt, double t0, double rdt);

/* ------------------------------------------------------------------ */
/* 3D acoustic CPML solver -- replaces a_iso_3d.c:368-417, 494-693     */
/* ------------------------------------------------------------------ */
typedef struct {
    size_t nx, ny, nz;        /* interior nodes (struct grid3d, structs.h:72-78) */
    size_t np;                /* CPML layers, 0 without absorbing boundary */
    double dx, dy, dz;        /* m */
    double dt;                /* s */
    double vmax;              /* abs_params.vmax (5000 in the reference) */
    double pml_order;         /* abs_params.pmlOrder (2) */
    double Rc;                /* abs_params.Rc */
    double kappa_max;         /* abs_params.kappa_max */
    int alpha_max;            /* abs_params.alpha_max: 0 -> alpha = 0 */
    double f0;                /* frequency of source 0 (alpha_max = pi*f0) */
    int storage;              /* SM_STORE_* */
    int flags;                /* SM_A3D_* */
} sm_a3d_params;

#define SM_A3D_NAIVE 1        /* two plain kernels per step (no fusion) */
#define SM_A3D_PAD64 4        /* fused kernels: device rows padded to a multiple
                                 of 64 cells; without
                                 the flag rows start on 128-byte lines when that
                                 adds at most 5% (FP32) / 10% (16-bit) cells,
                                 else they are padded to 8 cells */
#define SM_A3D_SROUND 2       /* BF16/FP16 storage: stochastic rounding of the
                                 stored values (deterministic pseudo-random
                                 bits from element index and step) instead
                                 of round-to-nearest-even */
#define SM_A3D_MAT16 8        /* BF16/FP16 storage without SM_A3D_SROUND: the
                                 material (rho, K*dt) is held as FP16 with
                                 one power-of-two scale per field (4 instead
                                 of 8 bytes per cell; values rounded once from
                                 FP64 to 11 significant bits; a field whose
                                 nonzero magnitudes span more than 2^28 is
                                 refused); tile variants only, rows_per_block
                                 4 or 8 (no counterpart in the reference) */

/* Build a solver from the padded model as read_model_a leaves it
 * (io_utils.c:1029-1153): vp and rho hold nx2*ny2*nz2 values with the
 * CPML cells already filled by edge replication.  Performs the set-up of
 * a_iso_3d.c:368-417 (including its Vp maximum over the first nx*nz
 * entries, 368-374), the CPML tables of 500-576, uploads the model and
 * zeroes the wavefields (494-498).  Returns SM_ERR_CFL where the reference
 * aborts (411-417). */
int sm_a3d_create(sm_ctx *ctx, const sm_a3d_params *p, const double *vp,
                  const double *rho, sm_a3d **out);
int sm_a3d_destroy(sm_a3d *s);
/* Point sources, a_iso_3d.c:689-693: after step it, while it < length[n],
 * P[idx[n]] += dt*fct[n][it]*1e6.  fct concatenated as for
 * sm_compute_src_fct.  Replaces any earlier set. */
int sm_a3d_set_sources(sm_a3d *s, size_t nsrc, const size_t *idx,
                       const size_t *length, const double *fct);
/* Receivers: the value of comp[r] (SM_VX, SM_VY, SM_VZ, SM_P) at padded
 * index idx[r] is recorded after every step, for up to max_steps steps
 * from the current step.  The host applies the record predicate
 * (sm_record_due) when writing traces (a_iso_3d.c:698-741). */
int sm_a3d_set_receivers(sm_a3d *s, size_t nrec, const size_t *idx,
                         const int *comp, size_t max_steps);
/* Advance nsteps time steps (velocity sweep, pressure sweep, source,
 * receivers; a_iso_3d.c:593-693).  Queued on the context stream. */
int sm_a3d_run(sm_a3d *s, size_t nsteps);
/* steps advanced since creation or the last reset */
int sm_a3d_current_step(const sm_a3d *s, size_t *it);
/* recorded values of steps [it0, it0+count) as out[count][nrec] (FP64);
 * synchronises the stream */
int sm_a3d_get_traces(sm_a3d *s, size_t it0, size_t count, double *out);
/* whole padded field (SM_P, SM_VX, SM_VY, SM_VZ) in the reference layout */
int sm_a3d_get_field(sm_a3d *s, int comp, double *out);
/* nx2*nz2 x-z plane at padded y index iy, i-major (a_iso_3d.c:726-734) */
int sm_a3d_get_slice_xz(sm_a3d *s, int comp, size_t iy, double *out);
/* zero wavefields, memory variables and recorded traces; step count -> 0 */
int sm_a3d_reset(sm_a3d *s);
/* Launch shape of the fused FP32/BF16/FP16 kernel: waves per block along y
 * (1, 2, 4 or 8) and number of x chunks marched in parallel.  Results do
 * not depend on it; the autotuner (seismod_amd/autotune.py) picks it. */
int sm_a3d_set_launch(sm_a3d *s, int rows_per_block, int x_chunks);
int sm_a3d_get_launch(const sm_a3d *s, int *rows_per_block, int *x_chunks);
/* Fused kernel variant (no counterpart in the reference): SM_A3D_CELL, one
 * cell per lane in z tiles of 63 owned cells, loads one plane ahead;
 * SM_A3D_CELL2, the same with loads two planes ahead; SM_A3D_TILE, z tiles
 * of 64 cells aligned to the device rows whose fields are staged plane by
 * plane through LDS in 16-byte chunks (every row read once per plane) and
 * whose new state leaves as 16-byte stores; SM_A3D_TILE_SPLIT, the same
 * with the CPML work of each direction compiled in only for tiles and
 * ranges of planes that touch that direction's absorbing layer;
 * SM_A3D_TILE_NT / SM_A3D_TILE_SPLIT_NT, the same two with non-temporal
 * state stores (rows_per_block 4 or 8 only).  Values 1, 3 and 4 are not
 * variants (SM_ERR_ARG).  The variants compute the same values bit for
 * bit; setting one resets the launch shape to its default. */
#define SM_A3D_CELL 0
#define SM_A3D_CELL2 2
#define SM_A3D_TILE 5
#define SM_A3D_TILE_SPLIT 6
#define SM_A3D_TILE_NT 7
#define SM_A3D_TILE_SPLIT_NT 8
int sm_a3d_set_variant(sm_a3d *s, int variant);
int sm_a3d_get_variant(const sm_a3d *s, int *variant);
/* Per-launch timing: when on, every step kernel launch is bracketed by HIP
 * events on the context stream; get_timing returns the summed kernel time
 * and the number of timed launches since timing was switched on. */
int sm_a3d_set_timing(sm_a3d *s, int on);
int sm_a3d_get_timing(const sm_a3d *s, double *ms_total, size_t *launches);
/* device bytes this solver holds, and the bytes sm_a3d_create would need
 * for parameters p (for admission control) */
int sm_a3d_device_bytes(const sm_a3d *s, size_t *bytes);
size_t sm_a3d_bytes_needed(const sm_a3d_params *p, size_t nrec, size_t max_steps);

/* ------------------------------------------------------------------ */
/* 2D elastic split-PML solver -- replaces e_iso.c:423-500, 565-670,    */
/* propagate.c:2511-2670, pml.c:39-1403, src.c:170-184                  */
/* ------------------------------------------------------------------ */
typedef struct {
    size_t nx, nz;            /* interior nodes (struct grid, np = 0) */
    size_t npml;              /* PML layers (>= 1; e_iso.c:346, Q: abs 0) */
    double h;                 /* dx == dz, m (e_iso.c:363-369) */
    double dt;                /* s */
    int storage;              /* SM_STORE_F32 or SM_STORE_F64 */
    int q1_compat;            /* non-zero: use l2m/la[nx*nx-1] in the
                                 lower-right corner of pml_t as pml.c:1044-1084
                                 does (only where nx <= nz; otherwise that
                                 read is out of bounds and the true corner
                                 cell nx*nz-1 is used) */
    double damping_scale;     /* 1.0 reproduces e_iso.c:490-500 */
    int flags;                /* SM_E2D_* */
} sm_e2d_params;

#define SM_E2D_NAIVE 1        /* FP32 with the reference-order two-kernel
                                 step instead of the fused one */

/* Build a solver from interior vp, vs, rho (nx*nz each, read_model_e,
 * io_utils.c:892-947): Lame constants and staggered averages
 * (e_iso.c:423-462), CFL check (476-479, SM_ERR_CFL), PML damping from
 * the mean Vp (482-500); zero fields (565-572). */
int sm_e2d_create(sm_ctx *ctx, const sm_e2d_params *p, const double *vp,
                  const double *vs, const double *rho, sm_e2d **out);
int sm_e2d_destroy(sm_e2d *s);
/* add_src (src.c:170-184): type[n] (SM_SX/SM_SZ/SM_BULK/SM_SXZ; the other
 * types of enum sourceType, 0..12, inject nothing; values outside it are
 * SM_ERR_ARG), interior index idx[n] = i*nz+j. */
int sm_e2d_set_sources(sm_e2d *s, size_t nsrc, const int *type,
                       const size_t *idx, const size_t *length,
                       const double *fct);
/* receivers on interior indices; comp SM_VX, SM_VZ, SM_TXX, SM_TZZ, SM_TXZ */
int sm_e2d_set_receivers(sm_e2d *s, size_t nrec, const size_t *idx,
                         const int *comp, size_t max_steps);
/* velocity phase, stress phase, source, receivers (e_iso.c:589-636) */
int sm_e2d_run(sm_e2d *s, size_t nsteps);
int sm_e2d_current_step(const sm_e2d *s, size_t *it);
int sm_e2d_get_traces(sm_e2d *s, size_t it0, size_t count, double *out);
/* interior field (nx*nz) as the reference's arrays hold it */
int sm_e2d_get_field(sm_e2d *s, int comp, double *out);
/* field on the padded (nx+2Np)x(nz+2Np) grid: interior values plus x+z of
 * the split pair in PML cells (what pml.c reads as a neighbour value) */
int sm_e2d_get_field_padded(sm_e2d *s, int comp, double *out);
int sm_e2d_reset(sm_e2d *s);
int sm_e2d_device_bytes(const sm_e2d *s, size_t *bytes);
/* 0 when sm_e2d_create would refuse the grid for its size: the fused FP32
 * solver needs its state and split fields, 96 bytes per padded cell,
 * below 2 GiB (about 4730^2 padded cells) */
size_t sm_e2d_bytes_needed(const sm_e2d_params *p, size_t nrec, size_t max_steps);

/* Launch shape of the fused FP32 step (no counterpart in the reference):
 * x rows each wavefront marches over, and steps per captured hipGraph
 * (0 or 1: plain launches).  Results do not depend on either. */
int sm_e2d_set_launch(sm_e2d *s, int rows_per_chunk, int graph_steps);
int sm_e2d_get_launch(const sm_e2d *s, int *rows_per_chunk, int *graph_steps);
/* Kernel form of the fused FP32 step (no counterpart in the reference):
 * SM_E2D_ROWS, each lane loads the rows of the next x row itself;
 * SM_E2D_STAGED, the rows every step needs arrive as 16-byte chunk loads
 * through LDS.  Results are identical. */
#define SM_E2D_ROWS 0
#define SM_E2D_STAGED 1
int sm_e2d_set_variant(sm_e2d *s, int variant);
int sm_e2d_get_variant(const sm_e2d *s, int *variant);
/* HIP-event timing of every sm_e2d_run (no counterpart in the reference) */
int sm_e2d_set_timing(sm_e2d *s, int on);
int sm_e2d_get_timing(const sm_e2d *s, double *ms_total, size_t *launches);

/* ------------------------------------------------------------------ */
/* Batches of independent 2D shots (no counterpart in the reference,   */
/* which runs one shot per process).  One launch per time step advances */
/* every shot that still has steps to run; long stretches are replayed  */
/* from captured hipGraphs.  Each shot's results are those of running   */
/* it alone with sm_e2d_run.                                            */
/* ------------------------------------------------------------------ */
typedef struct sm_e2d_batch sm_e2d_batch;
int sm_e2d_batch_create(sm_ctx *ctx, sm_e2d_batch **out);
int sm_e2d_batch_destroy(sm_e2d_batch *b);
/* steps per captured graph (0 or 1: plain launches; default 32) */
int sm_e2d_batch_set_graph_steps(sm_e2d_batch *b, int graph_steps);
/* advance shots[k] (FP32 fused solvers on the batch's device, each listed
 * once) by nsteps[k] steps; returns when all are done */
int sm_e2d_batch_run(sm_e2d_batch *b, size_t nshots, sm_e2d *const *shots,
                     const size_t *nsteps);
/* kernel form for the batch (SM_E2D_ROWS or SM_E2D_STAGED; the shots'
 * own settings apply to sm_e2d_run only) */
int sm_e2d_batch_set_variant(sm_e2d_batch *b, int variant);
int sm_e2d_batch_set_timing(sm_e2d_batch *b, int on);
int sm_e2d_batch_get_timing(const sm_e2d_batch *b, double *ms_total, size_t *launches);
/* Shot grouping of sm_e2d_batch_run (no counterpart in the reference): with
 * bytes > 0 the shots, longest run first, are packed into groups whose
 * wavefield, material and damping memory stays within bytes (a larger shot
 * is a group of its own), and each group runs all its steps before the
 * next starts, so its fields stay in the last-level cache between steps.
 * 0 (the default) advances every shot in every step launch.  Results do
 * not depend on it. */
int sm_e2d_batch_set_group_bytes(sm_e2d_batch *b, size_t bytes);
/* Allocation-free variant of sm_e2d_batch_run.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory
 * (serving loops and harnesses that forbid hipMalloc; it synchronises the
 * context's stream, so it is not for use under graph capture).  The same
 * kernels and results as sm_e2d_batch_run, with plain launches (no graph
 * replay, no timing); the shot table the kernels read lives in a
 * file-scope device array of 1 MiB per GPU (held until the process ends)
 * instead of memory the batch allocates; no hipMalloc or hipFree is
 * called.  A table larger than that (about
 * 250 000 blocks of four waves: some 60 M padded cells at one row per
 * chunk) is refused with SM_ERR_ARG.  Calls that overlap on one GPU, from
 * two host threads, race on that array; calls on one context are
 * serialised by the call itself. */
int sm_e2d_batch_run_noalloc(sm_e2d_batch *b, size_t nshots, sm_e2d *const *shots,
                             const size_t *nsteps);
/* Captured hipGraphs the batch holds (at most 16: the least recently used
 * one is destroyed to make room) and has instantiated since creation (no
 * counterpart in the reference). */
int sm_e2d_batch_get_graphs(const sm_e2d_batch *b, size_t *live, size_t *built);

/* ------------------------------------------------------------------ */
/* 2.5D elastic solver in cylindrical coordinates (e_cyl_iso) --        */
/* replaces e_cyl_iso.c:342-458, 526-535, 601-866, propagate.c:2706-4119 */
/* (update_*_cyl, update_*_cyl_n, compute_div_cyl, compute_curl_cyl),   */
/* pml.c:2107-2164 (compute_cpml_cyl), src.c:186-238                    */
/* ------------------------------------------------------------------ */
typedef struct {
    size_t nx, nz;            /* interior nodes along r and z (struct grid) */
    size_t np;                /* CPML layers: outer r side and both z ends
                                 (0: abs off, e_cyl_iso runs without layers) */
    double dr, dz;            /* m (g.dx, g.dz) */
    double dt;                /* s */
    int mode;                 /* azimuthal mode n (azimuthal_mode, >= 0) */
    double vmax, pml_order, Rc, kappa_max;   /* struct abs_params */
    int alpha_max;            /* abs_params.alpha_max: 0 -> alpha = 0 */
    double f0;                /* frequency of source 0 (alpha_max = pi*f0) */
    int storage;              /* SM_STORE_F32 or SM_STORE_F64 */
    int flags;                /* 0 */
} sm_cyl_params;

/* Build a solver from the padded model as read_model_e_cyl leaves it
 * (io_utils.c:949-1027): vp, vs, rho hold nx2*nz2 values, nx2 = nx+np,
 * nz2 = nz+2np, index i*nz2+j with row i at r = i*dr.  Computes the
 * staggered material (e_cyl_iso.c:342-415), refuses a non-finite value
 * (SM_ERR_ARG, 417-452) and a time step beyond 0.5(dr+dz)/(sqrt(2) Vp max)
 * (SM_ERR_CFL, 454-458), builds the CPML tables and zeroes the fields. */
int sm_cyl_create(sm_ctx *ctx, const sm_cyl_params *p, const double *vp, const double *vs,
                  const double *rho, sm_cyl **out);
int sm_cyl_destroy(sm_cyl *s);
/* Sources on padded indices.  SM_FR, SM_FT, SM_FZ and SM_BULK act as
 * forces after the velocity update (add_force_src_cyl, src.c:186-200:
 * v += dt * (1/rho) * fct); a SM_KURKJIAN source must be entry 0 and is
 * added to tau_rr and tau_zz after the stress update (src.c:231-238);
 * other types inject nothing.  fct concatenated as for sm_compute_src_fct
 * (or from sm_kurkjian_fct). */
int sm_cyl_set_sources(sm_cyl *s, size_t nsrc, const int *type, const size_t *idx,
                       const size_t *length, const double *fct);
/* Kurkjian source function (compute_kurkjian, src.c:202-229): lambda is the
 * Lame constant at the source cell, r its radial coordinate, A the peak
 * volume change. */
int sm_kurkjian_fct(double lambda, double dr, double dz, double r, int mode, double A, double f,
                    double dt, size_t length, double *fct);
/* receivers on padded indices; comp SM_VR, SM_VT, SM_VZ, SM_TRR, SM_TRT,
 * SM_TRZ, SM_TTT, SM_TTZ, SM_TZZ, SM_DIV, SM_CURL (e_cyl_iso.c:688-716) */
int sm_cyl_set_receivers(sm_cyl *s, size_t nrec, const size_t *idx, const int *comp,
                         size_t max_steps);
int sm_cyl_run(sm_cyl *s, size_t nsteps);
/* Per-step timing (no counterpart in the reference): when on, HIP events on
 * the solver's stream bracket the two update launches of every step;
 * totals accumulate until the next sm_cyl_set_timing. */
int sm_cyl_set_timing(sm_cyl *s, int on);
int sm_cyl_get_timing(const sm_cyl *s, double *ms_total, size_t *steps);
int sm_cyl_current_step(const sm_cyl *s, size_t *it);
int sm_cyl_get_traces(sm_cyl *s, size_t it0, size_t count, double *out);
/* padded nx2*nz2 field of comp (SM_VR ... SM_TTZ, SM_VZ, SM_TZZ) */
int sm_cyl_get_field(sm_cyl *s, int comp, double *out);
int sm_cyl_reset(sm_cyl *s);
int sm_cyl_device_bytes(const sm_cyl *s, size_t *bytes);
size_t sm_cyl_bytes_needed(const sm_cyl_params *p, size_t nrec, size_t max_steps);

/* ------------------------------------------------------------------ */
/* Viscoelastic SH waves in a 2D VTI medium (ve_vti_sh_pml) --          */
/* pseudo-spectral derivatives (LDS FFTs), CPML, fourth-order Runge-Kutta;*/
/* replaces ve_vti_sh_pml.c:475-700, propagate.c:540-591, 2280-2510     */
/* (init_fftw_data_ve_sh, propagateVE_SH_CPML, partial_*_sh), pml.c:    */
/* 1689-1757 and 1959-2106 (alloc_cpml_ve_sh, compute_cpml), src.c:     */
/* 145-168 (add_src_ve_sh_rk4)                                          */
/* ------------------------------------------------------------------ */
typedef struct sm_sh sm_sh;
/* flags of sm_sh_params and sm_ve_params (no counterpart in the reference):
 * take every derivative through hipFFT even where the hand-written
 * LDS transforms apply (axis lengths whose factors are 2, 3, 5, 7, 11, 13);
 * both give the same results to rounding */
#define SM_PSTD_HIPFFT 1
typedef struct {
    size_t nx, nz;            /* interior nodes (struct grid) */
    size_t np;                /* CPML layers on every side (nl); forced to 0
                                 when abs is 0 (io_utils.c:293) */
    double dx, dz;            /* m */
    double dt;                /* s */
    double vmax, pml_order, Rc, kappa_max;   /* struct abs_params */
    int alpha_max;            /* 0 -> alpha = 0, else alpha_max = pi*f0 */
    double f0;                /* frequency of source 0 */
    int abs;                  /* iwipe: 1 absorbing layers, 0 none */
    int storage;              /* SM_STORE_F64 or SM_STORE_F32 */
    int flags;                /* 0 or SM_PSTD_HIPFFT */
} sm_sh_params;

/* Build a solver from the padded model as read_model_ve_sh leaves it
 * (io_utils.c:1316-1429): mu = Vs0^2 rho 1e-9 (GPa), rho (kg/m^3), gamma,
 * Q and f hold nx2*nz2 values, nx2 = nx+2np, index i*nz2+j.  Computes the
 * staggered standard-linear-solid material (ve_vti_sh_pml.c:475-585), the
 * CPML tables of the three Runge-Kutta stage sets and the FFT plans, and
 * zeroes the fields.  Needs nx, nz >= 2 and rho, Q, f > 0 (SM_ERR_ARG). */
int sm_sh_create(sm_ctx *ctx, const sm_sh_params *p, const double *mu, const double *rho,
                 const double *gamma, const double *Q, const double *f, sm_sh **out);
int sm_sh_destroy(sm_sh *s);
/* Sources on padded indices: SM_SXY adds to tau_xy, SM_SYZ to tau_yz,
 * other types inject nothing (add_src_ve_sh_rk4).  fct holds each
 * function at half time steps, already divided by dx*dz, concatenated:
 * sm_compute_src_fct with dt/2 (compute_src_fct(.., dt, 0.5)) then the
 * 1/(dx dz) of normalize_src_fct (src.c:67-75); step it uses sample
 * 2 it + 1. */
int sm_sh_set_sources(sm_sh *s, size_t nsrc, const int *type, const size_t *idx,
                      const size_t *length, const double *fct);
/* receivers on padded indices; comp SM_VY, SM_TXY or SM_TYZ */
int sm_sh_set_receivers(sm_sh *s, size_t nrec, const size_t *idx, const int *comp,
                        size_t max_steps);
int sm_sh_run(sm_sh *s, size_t nsteps);
/* Per-step timing (no counterpart in the reference): HIP events around the
 * four Runge-Kutta stages of every step. */
int sm_sh_set_timing(sm_sh *s, int on);
int sm_sh_get_timing(const sm_sh *s, double *ms_total, size_t *steps);
int sm_sh_current_step(const sm_sh *s, size_t *it);
int sm_sh_get_traces(sm_sh *s, size_t it0, size_t count, double *out);
/* padded nx2*nz2 field: SM_VY, SM_TXY or SM_TYZ */
int sm_sh_get_field(sm_sh *s, int comp, double *out);
int sm_sh_reset(sm_sh *s);
int sm_sh_device_bytes(const sm_sh *s, size_t *bytes);
size_t sm_sh_bytes_needed(const sm_sh_params *p, size_t nrec, size_t max_steps);

/* ------------------------------------------------------------------ */
/* Viscoelastic P-SV waves in a 2D VTI medium (ve_vti_pml) --           */
/* pseudo-spectral derivatives (LDS FFTs), CPML, fourth-order Runge-Kutta,*/
/* L relaxation mechanisms for the quasi-dilatational mode and one for   */
/* shear; replaces ve_vti_pml.c:518-784, propagate.c:325-510 and         */
/* 1959-2278 (init_fftw_data_ve, propagateVE_CPML, partial_*), pml.c:    */
/* 1553-1631 and 1959-2106 (alloc_cpml_ve, compute_cpml), src.c:114-142  */
/* (add_src_ve_rk4)                                                      */
/* ------------------------------------------------------------------ */
typedef struct sm_ve sm_ve;
typedef struct {
    size_t nx, nz;            /* interior nodes (struct grid) */
    size_t np;                /* CPML layers on every side; forced to 0 when
                                 abs is 0 (io_utils.c:293) */
    double dx, dz;            /* m */
    double dt;                /* s */
    double vmax, pml_order, Rc, kappa_max;   /* struct abs_params */
    int alpha_max;            /* 0 -> alpha = 0, else alpha_max = pi*f0 */
    double f0;                /* frequency of source 0 */
    int abs;                  /* iwipe: 1 absorbing layers, 0 none */
    int L;                    /* quasi-dilatational relaxation mechanisms, >= 1 */
    int storage;              /* SM_STORE_F64 or SM_STORE_F32 */
    int flags;                /* 0 or SM_PSTD_HIPFFT */
} sm_ve_params;

/* Build a solver from the padded model as read_model_ve leaves it
 * (io_utils.c:1155-1314): K and mu in GPa, rho (kg/m^3), epsilon, delta,
 * Q2 and f2 hold nx2*nz2 values (nx2 = nx+2np, index i*nz2+j), Q1 and f1
 * L such arrays one after the other.  Computes the material of
 * ve_vti_pml.c:518-683, the CPML tables of the three Runge-Kutta stage sets
 * and the FFT plans, and zeroes the fields.  Needs nx, nz >= 2, rho, Q, f > 0
 * and a finite material (SM_ERR_ARG).  As in the reference, the e3 memory
 * variable does not feed back into tau_xz (c55_0_ij is never set there;
 * zero-filled memory) and the relaxation sum of tau_xx and tau_zz reads the
 * state at flat offset 6 nx2 nz2 + n + m (propagate.c:2190-2195). */
int sm_ve_create(sm_ctx *ctx, const sm_ve_params *p, const double *K, const double *mu,
                 const double *rho, const double *eps, const double *delta, const double *Q1,
                 const double *f1, const double *Q2, const double *f2, sm_ve **out);
int sm_ve_destroy(sm_ve *s);
/* The moduli ve_vti_pml writes with write_field_nc (ve_vti_pml.c:695-699),
 * in MPa on the padded grid, from the same inputs as sm_ve_create: c11,
 * c13, c33 and K0 at the nodes, c55 at (i+1/2, j+1/2).  Host only. */
int sm_ve_moduli(size_t nx2, size_t nz2, int L, const double *K, const double *mu,
                 const double *rho, const double *eps, const double *delta, const double *Q1,
                 const double *f1, const double *Q2, const double *f2, double *c11, double *c13,
                 double *c33, double *c55, double *K0);
/* Sources on padded indices: SM_SX adds to tau_xx, SM_SZ to tau_zz,
 * SM_BULK to both, SM_SXZ to tau_xz, other types inject nothing
 * (add_src_ve_rk4).  fct as for sm_sh_set_sources. */
int sm_ve_set_sources(sm_ve *s, size_t nsrc, const int *type, const size_t *idx,
                      const size_t *length, const double *fct);
/* receivers on padded indices; comp SM_VX, SM_VZ, SM_TXX, SM_TZZ or SM_TXZ */
int sm_ve_set_receivers(sm_ve *s, size_t nrec, const size_t *idx, const int *comp,
                        size_t max_steps);
int sm_ve_run(sm_ve *s, size_t nsteps);
/* Per-step timing (no counterpart in the reference): HIP events around the
 * four Runge-Kutta stages of every step. */
int sm_ve_set_timing(sm_ve *s, int on);
int sm_ve_get_timing(const sm_ve *s, double *ms_total, size_t *steps);
int sm_ve_current_step(const sm_ve *s, size_t *it);
int sm_ve_get_traces(sm_ve *s, size_t it0, size_t count, double *out);
/* padded nx2*nz2 field: SM_VX, SM_VZ, SM_TXX, SM_TZZ or SM_TXZ */
int sm_ve_get_field(sm_ve *s, int comp, double *out);
int sm_ve_reset(sm_ve *s);
int sm_ve_device_bytes(const sm_ve *s, size_t *bytes);
size_t sm_ve_bytes_needed(const sm_ve_params *p, size_t nrec, size_t max_steps);

/* ------------------------------------------------------------------ */
/* Poroviscoelastic waves in a 2D isotropic medium (pve_iso_pml) --     */
/* Biot solid and fluid motion (Carcione and Helle 1999), one relaxation */
/* mechanism for the fluid pressure, pseudo-spectral derivatives, CPML, */
/* fourth-order Runge-Kutta with the stiff viscous coupling solved      */
/* analytically at the start of every step; replaces pve_iso_pml.c:     */
/* 795-1145 and 1249-1328, propagate.c:49-285 and 884-1329             */
/* (init_fftw_data, propagateCPML), compute_div / compute_curl          */
/* (propagate.c:2672-2704), src.c:77-112 (add_src_rk4)                  */
/* ------------------------------------------------------------------ */
typedef struct sm_pve sm_pve;
typedef struct {
    size_t nx, nz;            /* interior nodes (struct grid) */
    size_t np;                /* CPML layers on every side; forced to 0 when
                                 abs is 0 (io_utils.c:293) */
    double dx, dz;            /* m */
    double dt;                /* s */
    double vmax, pml_order, Rc, kappa_max;   /* struct abs_params */
    int alpha_max;            /* 0 -> alpha = 0, else alpha_max = pi*f0 */
    double f0;                /* frequency of source 0 */
    int abs;                  /* iwipe: 1 absorbing layers, 0 none */
    int storage;              /* SM_STORE_F64 or SM_STORE_F32 */
    int flags;                /* SM_PSTD_HIPFFT, SM_PVE_PML_MEDIAN */
} sm_pve_params;
/* flag of sm_pve_params (pml_median 1, pve_iso_pml.c:833-845): smooth the
 * isotropic medium's node material (mu, E, M, alpha, varphi, tau_s, rho,
 * rho_f, eta, kappa, m) in the absorbing layers with a 3 x 3 median, ring
 * by ring (pml.c:2254-2329); not for the VTI medium (SM_ERR_ARG) */
#define SM_PVE_PML_MEDIAN 2

/* Build a solver from the padded model as read_model leaves it
 * (io_utils.c:403-625), nx2*nz2 values each (index i*nz2+j): K_m, K_s,
 * K_f and mu in GPa, phi, rho_s and rho_f in kg/m^3, tortuosity T, eta in
 * cP, kappa in mD, Q and the relaxation frequency fq in Hz.  Computes the
 * material and the analytic-step coefficients of pve_iso_pml.c:795-1145,
 * the CPML tables and the FFT plans, and zeroes the fields.  Needs nx,
 * nz >= 2, K_m, mu, eta >= 0 (mu = 0 gives a zero shear modulus at the
 * cell corners it touches, 1/(sum of 1/mu)), positive K_s, K_f, porosity,
 * densities, T, kappa, Q and fq, and a finite material (SM_ERR_ARG). */
int sm_pve_create(sm_ctx *ctx, const sm_pve_params *p, const double *K_m, const double *K_s,
                  const double *K_f, const double *phi, const double *mu, const double *rho_s,
                  const double *rho_f, const double *T, const double *eta, const double *kappa,
                  const double *Q, const double *fq, sm_pve **out);
/* A VTI Biot medium (pve_vti_pml; read_modelVTI, io_utils.c:627-800, and
 * pve_vti_pml.c:787-1145): the inputs of sm_pve_create with tortuosity and
 * permeability along x (T1, kappa1) and z (T3, kappa3) and Thomsen's
 * epsilon and delta of the frame; c33 = K_m + 4/3 mu, c55 = mu,
 * c11 = (1 + 2 epsilon) c33, c13 from delta, the Biot modulus and the
 * effective-stress coefficients alpha1, alpha3 of the anisotropic frame.
 * The same time loop and records (the stage update of propagate.c:
 * 1514-1958).  Holds two material arrays more than the isotropic solver
 * (sm_pve_bytes_needed counts them). */
int sm_pve_create_vti(sm_ctx *ctx, const sm_pve_params *p, const double *K_m, const double *K_s,
                      const double *K_f, const double *phi, const double *mu, const double *rho_s,
                      const double *rho_f, const double *T1, const double *T3, const double *eta,
                      const double *kappa1, const double *kappa3, const double *Q, const double *fq,
                      const double *epsilon, const double *delta, sm_pve **out);
int sm_pve_destroy(sm_pve *s);
/* Sources on padded indices (add_src_rk4): SM_SF adds to p, SM_SX to
 * tau_xx, SM_SZ to tau_zz, SM_BULK to p, tau_xx and tau_zz, SM_BULK_S to
 * tau_xx and tau_zz, SM_SXZ to tau_xz; fct as for sm_sh_set_sources. */
int sm_pve_set_sources(sm_pve *s, size_t nsrc, const int *type, const size_t *idx,
                       const size_t *length, const double *fct);
/* receivers on padded indices: SM_VX, SM_VZ, SM_QX, SM_QZ, SM_TXX, SM_TZZ,
 * SM_TXZ, SM_P, or SM_DIV / SM_CURL (D_x^- v_x + D_z^- v_z and
 * D_x^+ v_z - D_z^+ v_x of the last Runge-Kutta stage's input, without the
 * layer corrections, as compute_div and compute_curl take them) */
int sm_pve_set_receivers(sm_pve *s, size_t nrec, const size_t *idx, const int *comp,
                         size_t max_steps);
int sm_pve_run(sm_pve *s, size_t nsteps);
/* Kinetic energy of a region after every step, from the current step on
 * (saveEnergy, pve_iso_pml.c:1106-1119 and 1391-1404): the sum over padded
 * cells i1 <= i <= i2, j1 <= j <= j2 of 0.5 rho11 (v_x^2 + v_z^2) +
 * rho12 (v_x q_x + v_z q_z) + 0.5 rho22 (q_x^2 + q_z^2), rho11, rho12 and
 * rho22 given per region cell (i outer, (i2-i1+1)(j2-j1+1) values each);
 * accumulated in FP64 in a fixed order (results repeat bit for bit).
 * max_steps 0 stops recording. */
int sm_pve_set_energy(sm_pve *s, size_t i1, size_t i2, size_t j1, size_t j2, const double *rho11,
                      const double *rho12, const double *rho22, size_t max_steps);
/* The same with different coefficients along x and z (pve_vti_pml.c:
 * 1106-1121 and 1334-1345): the sum of 0.5 ax v_x^2 + bx v_x q_x +
 * 0.5 cx q_x^2 + 0.5 az v_z^2 + bz v_z q_z + 0.5 cz q_z^2.  Adds 48 bytes
 * per region cell and 8 per step (so does sm_pve_set_energy). */
int sm_pve_set_energy_axes(sm_pve *s, size_t i1, size_t i2, size_t j1, size_t j2, const double *ax,
                           const double *bx, const double *cx, const double *az, const double *bz, const double *cz,
                           size_t max_steps);
/* count values from row it0 of the energy record */
int sm_pve_get_energy(sm_pve *s, size_t it0, size_t count, double *out);
/* Checkpoints (no counterpart API: the reference's save_checkpoint and
 * read_checkpoint2, io_utils.c:2355-2811, fwrite its structs): the state a
 * run continues from -- step count, current fields, CPML memory variables,
 * the traces and energy recorded so far -- as one host blob of
 * sm_pve_state_bytes bytes.  sm_pve_set_state takes a blob from a solver of
 * the same grid, storage, receivers and energy region (SM_ERR_ARG
 * otherwise); the next sm_pve_run continues the saved run bit for bit. */
int sm_pve_state_bytes(const sm_pve *s, size_t *bytes);
int sm_pve_get_state(sm_pve *s, void *buf, size_t bytes);
int sm_pve_set_state(sm_pve *s, const void *buf, size_t bytes);
int sm_pve_set_timing(sm_pve *s, int on);
int sm_pve_get_timing(const sm_pve *s, double *ms_total, size_t *steps);
int sm_pve_current_step(const sm_pve *s, size_t *it);
int sm_pve_get_traces(sm_pve *s, size_t it0, size_t count, double *out);
/* padded nx2*nz2 field of any receiver component */
int sm_pve_get_field(sm_pve *s, int comp, double *out);
int sm_pve_reset(sm_pve *s);
int sm_pve_device_bytes(const sm_pve *s, size_t *bytes);
size_t sm_pve_bytes_needed(const sm_pve_params *p, size_t nrec, size_t max_steps);

#ifdef __cplusplus
}
#endif
#endif /* SEISMOD_HIP_H */
