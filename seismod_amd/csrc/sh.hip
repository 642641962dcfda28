// sh.hip -- viscoelastic SH waves in a 2D VTI medium (ve_vti_sh_pml):
// pseudo-spectral derivatives on a staggered grid, CPML, fourth-order
// Runge-Kutta in time, one standard-linear-solid relaxation per stress.
//
// The reference (ve_vti_sh_pml.c:637-697, propagate.c:2280-2462) evaluates
// four derivative stages per step; each stage takes four FFT derivatives
// (D_x^- tau_xy, D_z^- tau_yz, D_z^+ v_y, D_x^+ v_y), corrects them in the
// CPML layers and forms the pointwise right-hand sides, which the
// Runge-Kutta combination then folds into the state.  Here a stage is the
// four derivatives (pstd.hpp: one LDS transform launch per axis, or
// hipFFT plans and a multiplier pass for other lengths) followed by ONE
// fused kernel that applies the layer
// corrections and memory-variable updates, the right-hand sides and the
// Runge-Kutta update of every state value.  The step-start copies of the
// state are replaced by rotating three state buffers (w, ws, wt), and the
// memory-variable copies are folded into the layer updates.  Results match
// the FP64 restatement to FFT rounding (the FFT libraries differ), not bit
// for bit.
//
// Deviations (undefined reads in the reference, given their natural
// values): c44_0 of the last x row uses that row's filled eta2; the tables
// of the right-side dx_txy loop at nl = np (one past their end) take their
// own formula there (k = 1, c = 0); dx_vy, indexed like a z layer, gets an
// array large enough for nx2 columns.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pstd.hpp"

namespace sm {
namespace {

using pstd::Layer;

enum { W_TXY, W_TYZ, W_VY, W_E1, W_E2, W_NF };
enum { M_C44, M_C66, M_C440, M_C660, M_TS1, M_TS2, M_ETA1, M_ETA2, M_IRHO, M_NMAT };
// layer sets (memory variables) in stage order
enum { L_DX_TXY, L_DZ_TYZ, L_DZ_VY, L_DX_VY, L_NSET };

template <typename T>
struct ShStage {
    const T *d;          // D_x^- tau_xy, D_z^- tau_yz, D_z^+ v_y, D_x^+ v_y (N each)
    const T *mat;        // M_NMAT x N
    const T *x;          // stage input state
    const T *ws;         // state at the start of the step
    T *w, *wt;
    T a, b;              // wt = ws + a delta (set_t); w = (acc ? w : ws) + b delta
    int set_t, acc;
    int nx2, nz2, np;
    unsigned N;
    Layer<T> L[L_NSET];
    pstd::StageSrc<T> src;
};

// add_src_ve_sh_rk4 (src.c:144-168) for source s: Sxy to tau_xy, Syz to
// tau_yz, other types nothing
template <typename T>
__device__ __forceinline__ void sh_add_src(T *w, T *wt, size_t N, int s, int nsrc, const int64_t *tab, const T *fct,
                                           T dtw, T dtt, int64_t it)
{
    const int64_t k = 2 * it + 1;
    const int64_t typ = tab[s], cell = tab[nsrc + s], off = tab[2 * nsrc + s], len = tab[3 * nsrc + s];
    if (k >= len) return;
    const size_t q = typ == SM_SXY ? 0 : typ == SM_SYZ ? N : ~(size_t)0;
    if (q == ~(size_t)0) return;
    const T v = fct[off + k];
    if (dtt != T(0)) wt[q + cell] += dtt * v;
    w[q + cell] += dtw * v;
}

// propagate.c:2280-2462 from the four derivatives on, then eq. (79)-(83),
// then the sources on this cell
template <typename T>
__global__ __launch_bounds__(256) void sh_stage(const ShStage<T> a)
{
    const unsigned n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= a.N) return;
    const unsigned N = a.N;
    const int i = (int)(n / (unsigned)a.nz2), j = (int)(n % (unsigned)a.nz2);
    const T d0 = pstd::layer_fix(a.L[L_DX_TXY], a.d[n], i, j, a.nx2, a.nz2, a.np);
    const T d1 = pstd::layer_fix(a.L[L_DZ_TYZ], a.d[N + n], i, j, a.nx2, a.nz2, a.np);
    const T d2 = pstd::layer_fix(a.L[L_DZ_VY], a.d[2 * N + n], i, j, a.nx2, a.nz2, a.np);
    const T d3 = pstd::layer_fix(a.L[L_DX_VY], a.d[3 * N + n], i, j, a.nx2, a.nz2, a.np);
    const T *m = a.mat;
    // v_y on the interior nodes only (propagate.c:2364-2374); elsewhere its
    // rate stays 0
    const bool inner = i >= 1 && i < a.nx2 - 1 && j >= 1 && j < a.nz2 - 1;
    T dl[W_NF];
    dl[W_VY] = inner ? m[M_IRHO * N + n] * (d0 + d1) : T(0);
    // e' = (1/tau_s)((1 - 1/eta) D v - e);  tau' = c D v + c_0 e
    const T e1 = a.x[W_E1 * N + n], e2 = a.x[W_E2 * N + n];
    dl[W_E1] = m[M_TS1 * N + n] * (m[M_ETA1 * N + n] * d2 - e1);
    dl[W_TYZ] = m[M_C44 * N + n] * d2 + m[M_C440 * N + n] * e1;
    dl[W_E2] = m[M_TS2 * N + n] * (m[M_ETA2 * N + n] * d3 - e2);
    dl[W_TXY] = m[M_C66 * N + n] * d3 + m[M_C660 * N + n] * e2;
#pragma unroll
    for (int c = 0; c < W_NF; ++c) {
        const size_t q = (size_t)c * N + n;
        const T s = a.ws[q];
        if (a.set_t) a.wt[q] = s + a.a * dl[c];
        a.w[q] = (a.acc ? a.w[q] : s) + a.b * dl[c];
    }
    for (int s = 0; s < a.src.n; ++s)
        if (a.src.cell[s] == (int)n)
            sh_add_src(a.w, a.wt, N, s, a.src.nsrc, a.src.tab, a.src.fct, a.src.dtw, a.src.dtt, a.src.it);
}

// all sources in order on one thread (more sources than the stage kernel
// takes)
template <typename T>
__global__ void sh_src(T *w, T *wt, size_t N, int nsrc, const int64_t *tab, const T *fct, T dtw, T dtt,
                       int64_t it)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int s = 0; s < nsrc; ++s) sh_add_src(w, wt, N, s, nsrc, tab, fct, dtw, dtt, it);
}

template <typename T>
__global__ void sh_gather(const T *w, size_t N, int nrec, const int64_t *ridx, const int *rcomp, T *trace,
                          int64_t row)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrec) return;
    const int c = rcomp[r];
    const size_t q = c == SM_TXY ? 0 : c == SM_TYZ ? N : 2 * N;
    trace[row * nrec + r] = w[q + ridx[r]];
}

}  // namespace
}  // namespace sm

using namespace sm;

struct sm_sh {
    sm_ctx *ctx = nullptr;
    sm_sh_params p{};
    pstd::Grid g;
    size_t N = 0, elt = 8;
    // three state buffers of 5 N (w, ws, wt by rotation), four derivative
    // fields, material (9 N)
    sm_dbuf state, der, mat;
    int iw = 0, iws = 1, iwt = 2;
    sm_dbuf mem;                    // 3 copies x (dx_txy, dx_vy, dz_tyz, dz_vy)
    size_t mem_dx = 0, mem_dxv = 0, mem_dz = 0;   // entries of dx_txy, dx_vy, dz_* per copy
    sm_dbuf src, rcv, trace;
    size_t nsrc = 0, nrec = 0, max_steps = 0, rcv_step0 = 0, step = 0;
    size_t src_cells[pstd::SRC_FUSED] = {};   // cells of the sources the stage kernel injects
    bool timing = false;
    double timed_ms = 0.0;
    size_t timed_steps = 0;
};

namespace {

void release(sm_sh *s)
{
    pstd::release(s->g);
    for (sm_dbuf *b : {&s->state, &s->der, &s->mat, &s->mem, &s->src, &s->rcv, &s->trace}) dfree(*b);
}

size_t mem_entries(size_t np, size_t nx2, size_t nz2, size_t *dx, size_t *dxv, size_t *dz)
{
    *dx = (2 * np + 1) * nz2;
    *dxv = (2 * np + 1) * std::max(nx2, nz2);
    *dz = (2 * np + 1) * nx2;
    return 3 * (*dx + *dxv + 2 * *dz);
}

template <typename T> T *STATE(sm_sh *s, int b) { return (T *)s->state.ptr + (size_t)b * W_NF * s->N; }
template <typename T> T *MEM(sm_sh *s, int copy, int set)   // set 0 dx_txy, 1 dx_vy, 2 dz_tyz, 3 dz_vy
{
    const size_t per = s->mem_dx + s->mem_dxv + 2 * s->mem_dz;
    const size_t off[4] = {0, s->mem_dx, s->mem_dx + s->mem_dxv, s->mem_dx + s->mem_dxv + s->mem_dz};
    return (T *)s->mem.ptr + copy * per + off[set];
}

// one propagation (stage tables `tabs`) of input x, then the Runge-Kutta
// update of the fused kernel
template <typename T>
int stage(sm_sh *s, int tabs, bool last, const T *x, T a, T b, int set_t, int acc, T dtw, T dtt, int64_t it)
{
    hipStream_t st = s->ctx->stream;
    const size_t N = s->N;
    const int np = (int)s->g.np;
    T *d = (T *)s->der.ptr;
    int rc;
    // D_x^- tau_xy, D_x^+ v_y; D_z^- tau_yz, D_z^+ v_y (field, slot, forward)
    const pstd::DJob xj[2] = {{W_TXY, L_DX_TXY, false}, {W_VY, L_DX_VY, true}};
    const pstd::DJob zj[2] = {{W_TYZ, L_DZ_TYZ, false}, {W_VY, L_DZ_VY, true}};
    if ((rc = pstd::deriv_xz<T>(s->g, x, d, xj, 2, zj, 2)) != SM_OK) return rc;
    ShStage<T> k;
    k.d = d; k.mat = (const T *)s->mat.ptr; k.x = x; k.ws = STATE<T>(s, s->iws);
    k.w = STATE<T>(s, s->iw); k.wt = STATE<T>(s, s->iwt);
    k.a = a; k.b = b; k.set_t = set_t; k.acc = acc;
    k.nx2 = s->g.nx2; k.nz2 = s->g.nz2; k.np = np; k.N = (unsigned)N;
    k.src.n = 0;
    if (s->nsrc && s->nsrc <= (size_t)pstd::SRC_FUSED) {
        k.src.n = (int)s->nsrc;
        for (size_t q = 0; q < s->nsrc; ++q) k.src.cell[q] = (int)s->src_cells[q];
        k.src.tab = (const int64_t *)s->src.ptr;
        k.src.fct = (const T *)(k.src.tab + 4 * s->nsrc);
        k.src.nsrc = (int)s->nsrc;
        k.src.dtw = dtw; k.src.dtt = dtt; k.src.it = it;
    }
    // copies of the reference: m2 <- m1 over (2np+1) rows of nz2 (dx_*) or
    // nx2 (dz_*) entries after every propagation (propagate.c:2451-2461);
    // m0 <- m1 over 2np rows at the start of a step (ve_vti_sh_pml.c:648-655)
    const size_t r2x = (size_t)(2 * np + 1) * s->g.nz2, r2z = (size_t)(2 * np + 1) * s->g.nx2;
    const size_t r0x = last ? (size_t)2 * np * s->g.nz2 : 0, r0z = last ? (size_t)2 * np * s->g.nx2 : 0;
    using namespace pstd;
    // dx_txy: left rows nl < np, right rows nl <= np, node tables
    k.L[L_DX_TXY] = layer<T>(s->g, tabs, MEM<T>(s, 0, 0), MEM<T>(s, 1, 0), MEM<T>(s, 2, 0), 0, np, 0, T_B, K_K, 0,
                             T_B, K_K, r2x, r0x);
    // dz_tyz: top and bottom rows nl < np
    k.L[L_DZ_TYZ] = layer<T>(s->g, tabs, MEM<T>(s, 0, 2), MEM<T>(s, 1, 2), MEM<T>(s, 2, 2), 1, np - 1, 1, T_B,
                             K_K, 1, T_B, K_K, r2z, r0z);
    // dz_vy: top inner half nodes, bottom outer half nodes nl <= np
    k.L[L_DZ_VY] = layer<T>(s->g, tabs, MEM<T>(s, 0, 3), MEM<T>(s, 1, 3), MEM<T>(s, 2, 3), 1, np, 1, T_BH,
                            K_KH, 1, T_BHE, K_KHE, r2z, r0z);
    // dx_vy: indexed like a z layer over nx2 columns, the x inner-half
    // tables on top and the z outer-half tables below, copied over nz2
    // columns (propagate.c:2410-2425, 2451-2455)
    k.L[L_DX_VY] = layer<T>(s->g, tabs, MEM<T>(s, 0, 1), MEM<T>(s, 1, 1), MEM<T>(s, 2, 1), 1, np, 0, T_BH,
                            K_KH, 1, T_BHE, K_KHE, r2x, r0x);
    sh_stage<T><<<pstd::blocks(N), 256, 0, st>>>(k);
    return SM_OK;
}

template <typename T>
int run(sm_sh *s, size_t nsteps)
{
    hipStream_t st = s->ctx->stream;
    const size_t N = s->N;
    const T dt = (T)s->p.dt;
    const int64_t *stab = (const int64_t *)s->src.ptr;
    const T *sfct = stab ? (const T *)(stab + 4 * s->nsrc) : nullptr;
    const int64_t *ridx = (const int64_t *)s->rcv.ptr;
    const int *rcomp = ridx ? (const int *)(ridx + s->nrec) : nullptr;
    std::vector<hipEvent_t> ev;
    struct Guard {
        std::vector<hipEvent_t> &e;
        ~Guard() { for (hipEvent_t x : e) if (x) (void)hipEventDestroy(x); }
    } guard{ev};
    if (s->timing) {
        ev.assign(2 * nsteps, nullptr);
        for (auto &e : ev) SM_HIP(hipEventCreate(&e));
    }
    // eq. (80)-(83) and (79) of the reference's scheme: tables, wt = ws + a
    // delta, w = (acc ? w : ws) + b delta, source weights (w, wt)
    const T h6 = dt / T(6), h3 = dt / T(3), h2 = dt / T(2);
    struct Stage { int tabs; T a; T b; int acc; T sw; T st; };
    const Stage stages[4] = {{0, h2, h6, 0, h6, h2}, {1, h2, h3, 1, h3, h2},
                             {1, dt, h3, 1, h3, dt}, {2, T(0), h6, 1, h6, T(0)}};
    for (size_t k = 0; k < nsteps; ++k) {
        const int64_t it = (int64_t)(s->step + k);
        if (s->timing) SM_HIP(hipEventRecord(ev[2 * k], st));
        // Ws = W: the buffer holding w becomes ws; stage 1 rewrites all of w
        std::swap(s->iw, s->iws);
        for (int q = 0; q < 4; ++q) {
            const T *x = q == 0 ? STATE<T>(s, s->iws) : STATE<T>(s, s->iwt);
            int rc = stage<T>(s, stages[q].tabs, q == 3, x, stages[q].a, stages[q].b, q < 3 ? 1 : 0,
                              stages[q].acc, stages[q].sw, stages[q].st, it);
            if (rc != SM_OK) return rc;
            if (s->nsrc > (size_t)pstd::SRC_FUSED)
                sh_src<T><<<1, 64, 0, st>>>(STATE<T>(s, s->iw), STATE<T>(s, s->iwt), N, (int)s->nsrc, stab, sfct,
                                            stages[q].sw, stages[q].st, it);
        }
        if (s->timing) SM_HIP(hipEventRecord(ev[2 * k + 1], st));
        if (s->nrec && (size_t)it >= s->rcv_step0 && (size_t)it - s->rcv_step0 < s->max_steps)
            sh_gather<T><<<pstd::blocks(s->nrec), 256, 0, st>>>(STATE<T>(s, s->iw), N, (int)s->nrec, ridx, rcomp,
                                                                (T *)s->trace.ptr,
                                                                (int64_t)((size_t)it - s->rcv_step0));
    }
    SM_HIP(hipGetLastError());
    s->step += nsteps;
    if (s->timing) {
        SM_HIP(hipStreamSynchronize(st));
        for (size_t k = 0; k < nsteps; ++k) {
            float ms = 0.f;
            SM_HIP(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
            s->timed_ms += ms;
            s->timed_steps += 1;
        }
    }
    return SM_OK;
}

int zero_state(sm_sh *s)
{
    hipStream_t st = s->ctx->stream;
    SM_HIP(hipMemsetAsync(s->state.ptr, 0, s->state.bytes, st));
    SM_HIP(hipMemsetAsync(s->mem.ptr, 0, s->mem.bytes, st));
    if (s->trace.ptr) SM_HIP(hipMemsetAsync(s->trace.ptr, 0, s->trace.bytes, st));
    SM_HIP(hipStreamSynchronize(st));
    s->iw = 0; s->iws = 1; s->iwt = 2;
    return SM_OK;
}

template <typename T>
int read_back(sm_sh *s, const void *dev, size_t n, double *out)
{
    std::vector<T> tmp(n);
    SM_HIP(hipMemcpyAsync(tmp.data(), dev, n * sizeof(T), hipMemcpyDeviceToHost, s->ctx->stream));
    SM_HIP(hipStreamSynchronize(s->ctx->stream));
    for (size_t i = 0; i < n; ++i) out[i] = (double)tmp[i];
    return SM_OK;
}

// the staggered material of ve_vti_sh_pml.c:475-585 (the last row's c44_0
// uses the filled eta2, see the file comment)
bool material(size_t nx2, size_t nz2, const double *mu, const double *rho, const double *gamma,
              const double *Q, const double *f, std::vector<double> &all)
{
    const size_t N = nx2 * nz2;
    const double pi = 4.0 * std::atan(1.0);
    std::vector<double> m[M_NMAT];
    for (auto &a : m) a.assign(N, 0.0);
    std::vector<double> mu_(N), rho_(N);
    for (size_t n = 0; n < N; ++n) { mu_[n] = mu[n] * 1000.; rho_[n] = rho[n] * 1.e-6; }
    for (size_t i = 0; i + 1 < nx2; ++i)
        for (size_t j = 0; j < nz2; ++j) {
            const size_t ij = i * nz2 + j, iij = (i + 1) * nz2 + j;
            const double q = 0.5 * (Q[ij] + Q[iij]), fr = 0.5 * (f[ij] + f[iij]);
            const double tau_e = (std::sqrt(q * q + 1) + 1.) / (2. * pi * fr * q);
            m[M_TS2][ij] = (std::sqrt(q * q + 1) - 1.) / (2. * pi * fr * q);
            m[M_ETA2][ij] = m[M_TS2][ij] / tau_e;
            const double c44 = 0.5 * (mu_[ij] + mu_[iij]);
            const double g = 0.5 * (gamma[ij] + gamma[iij]);
            m[M_C66][ij] = 2. * c44 * g + c44;
            m[M_C660][ij] = m[M_C66][ij] * m[M_ETA2][ij];
        }
    for (size_t j = 0; j < nz2; ++j) {
        const size_t ij = (nx2 - 2) * nz2 + j, iij = (nx2 - 1) * nz2 + j;
        for (int q : {M_C66, M_C660, M_TS2, M_ETA2}) m[q][iij] = m[q][ij];
    }
    for (size_t i = 0; i < nx2; ++i)
        for (size_t j = 0; j + 1 < nz2; ++j) {
            const size_t ij = i * nz2 + j, ijj = ij + 1;
            const double q = 0.5 * (Q[ij] + Q[ijj]), fr = 0.5 * (f[ij] + f[ijj]);
            const double tau_e = (std::sqrt(q * q + 1) + 1.) / (2. * pi * fr * q);
            m[M_TS1][ij] = (std::sqrt(q * q + 1) - 1.) / (2. * pi * fr * q);
            m[M_ETA1][ij] = m[M_TS1][ij] / tau_e;
            m[M_C44][ij] = 0.5 * (mu_[ij] + mu_[ijj]);
            m[M_C440][ij] = m[M_C44][ij] * m[M_ETA2][ij];
        }
    for (size_t i = 0; i < nx2; ++i) {
        const size_t ij = i * nz2 + nz2 - 2, ijj = i * nz2 + nz2 - 1;
        for (int q : {M_C44, M_C440, M_TS1, M_ETA1}) m[q][ijj] = m[q][ij];
    }
    for (size_t n = 0; n < N; ++n) {
        m[M_IRHO][n] = 1. / rho_[n];
        m[M_ETA1][n] = 1. - 1. / m[M_ETA1][n];
        m[M_ETA2][n] = 1. - 1. / m[M_ETA2][n];
        m[M_TS1][n] = 1. / m[M_TS1][n];
        m[M_TS2][n] = 1. / m[M_TS2][n];
    }
    all.clear();
    all.reserve(M_NMAT * N);
    for (auto &a : m) all.insert(all.end(), a.begin(), a.end());
    for (double v : all)
        if (!std::isfinite(v)) return false;
    return true;
}

}  // namespace

extern "C" {

size_t sm_sh_bytes_needed(const sm_sh_params *p, size_t nrec, size_t max_steps)
{
    if (!p || (p->storage != SM_STORE_F64 && p->storage != SM_STORE_F32)) return 0;
    const size_t e = p->storage == SM_STORE_F64 ? 8 : 4;
    const size_t np = p->abs ? p->np : 0;
    const size_t nx2 = p->nx + 2 * np, nz2 = p->nz + 2 * np, N = nx2 * nz2;
    size_t a, b, c;
    const size_t mem = mem_entries(np, nx2, nz2, &a, &b, &c);
    return e * (3 * W_NF * N + 4 * N + M_NMAT * N + mem) + pstd::grid_bytes(nx2, nz2, np, e) +
           e * nrec * max_steps + 12 * nrec;
}

int sm_sh_create(sm_ctx *ctx, const sm_sh_params *pp, const double *mu, const double *rho,
                 const double *gamma, const double *Q, const double *f, sm_sh **out)
{
    SM_CHECK(ctx && pp && mu && rho && gamma && Q && f && out, SM_ERR_ARG, "sm_sh_create: NULL argument");
    sm_sh_params p = *pp;
    SM_CHECK(p.storage == SM_STORE_F64 || p.storage == SM_STORE_F32, SM_ERR_ARG,
             "sm_sh_create: storage must be SM_STORE_F64 or SM_STORE_F32");
    SM_CHECK(p.nx >= 2 && p.nz >= 2, SM_ERR_ARG, "sm_sh_create: the model needs at least 2 x 2 nodes");
    SM_CHECK(p.dx > 0 && p.dz > 0 && p.dt > 0, SM_ERR_ARG, "sm_sh_create: spacings and dt must be > 0");
    SM_CHECK((p.flags & ~SM_PSTD_HIPFFT) == 0, SM_ERR_ARG, "sm_sh_create: unknown flags %#x", p.flags);
    if (!p.abs) p.np = 0;   // io_utils.c:293: no absorbing layers without abs
    const size_t nx2 = p.nx + 2 * p.np, nz2 = p.nz + 2 * p.np, N = nx2 * nz2;
    SM_CHECK(nx2 < (1u << 20) && nz2 < (1u << 20) && N < ((size_t)1 << 30), SM_ERR_ARG,
             "sm_sh_create: grid too large");
    for (size_t n = 0; n < N; ++n)
        SM_CHECK(rho[n] > 0 && std::isfinite(mu[n]) && std::isfinite(gamma[n]) && Q[n] > 0 && f[n] > 0,
                 SM_ERR_ARG, "sm_sh_create: cell %zu needs rho, Q, f > 0 and finite mu, gamma", n);
    std::vector<double> mat;
    SM_CHECK(material(nx2, nz2, mu, rho, gamma, Q, f, mat), SM_ERR_ARG, "sm_sh_create: material not finite");
    SM_CHECK(set_device(ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");

    sm_sh *s = new sm_sh;
    s->ctx = ctx;
    s->p = p;
    s->N = N;
    s->elt = p.storage == SM_STORE_F64 ? 8 : 4;
    const size_t e = s->elt;
    const size_t mem = mem_entries(p.np, nx2, nz2, &s->mem_dx, &s->mem_dxv, &s->mem_dz);
    int rc = dalloc(s->state, 3 * W_NF * N * e);
    if (rc == SM_OK) rc = dalloc(s->der, 4 * N * e);
    if (rc == SM_OK) rc = dalloc(s->mat, M_NMAT * N * e);
    if (rc == SM_OK) rc = dalloc(s->mem, mem * e);
    pstd::Cpml cp;
    cp.abs = p.abs; cp.vmax = p.vmax; cp.pml_order = p.pml_order; cp.Rc = p.Rc; cp.kappa_max = p.kappa_max;
    cp.alpha_max = p.alpha_max;
    if (rc == SM_OK)
        rc = pstd::init(s->g, ctx, nx2, nz2, p.np, e, p.dx, p.dz, p.dt, p.f0, cp, !(p.flags & SM_PSTD_HIPFFT));
    if (rc == SM_OK) rc = pstd::upload_any(e, s->mat.ptr, mat, ctx->stream);
    if (rc == SM_OK) rc = zero_state(s);
    if (rc != SM_OK) { release(s); delete s; return rc; }
    *out = s;
    return SM_OK;
}

int sm_sh_destroy(sm_sh *s)
{
    if (!s) return SM_OK;
    (void)set_device(s->ctx);
    (void)hipStreamSynchronize(s->ctx->stream);
    release(s);
    delete s;
    return SM_OK;
}

int sm_sh_set_sources(sm_sh *s, size_t nsrc, const int *type, const size_t *idx, const size_t *length,
                      const double *fct)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_sh_set_sources: NULL solver");
    SM_CHECK(nsrc == 0 || (type && idx && length && fct), SM_ERR_ARG, "sm_sh_set_sources: NULL array");
    SM_CHECK(nsrc < (1u << 20), SM_ERR_ARG, "sm_sh_set_sources: too many sources");
    size_t total = 0;
    for (size_t n = 0; n < nsrc; ++n) {
        SM_CHECK(idx[n] < s->N, SM_ERR_ARG, "source %zu: index %zu outside the grid", n, idx[n]);
        SM_CHECK(type[n] >= SM_SX && type[n] <= SM_KURKJIAN, SM_ERR_ARG, "source %zu: unknown type %d", n,
                 type[n]);
        total += length[n];
    }
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    dfree(s->src);
    s->nsrc = nsrc;
    if (nsrc == 0) return SM_OK;
    const size_t tab = 4 * nsrc * sizeof(int64_t);
    std::vector<unsigned char> h(tab + total * s->elt);
    int64_t *t = (int64_t *)h.data();
    size_t off = 0;
    for (size_t n = 0; n < nsrc; ++n) {
        if (n < (size_t)pstd::SRC_FUSED) s->src_cells[n] = idx[n];
        t[n] = type[n]; t[nsrc + n] = (int64_t)idx[n]; t[2 * nsrc + n] = (int64_t)off;
        t[3 * nsrc + n] = (int64_t)length[n];
        for (size_t k = 0; k < length[n]; ++k) {
            if (s->elt == 8) ((double *)(h.data() + tab))[off + k] = fct[off + k];
            else ((float *)(h.data() + tab))[off + k] = (float)fct[off + k];
        }
        off += length[n];
    }
    int rc = dalloc(s->src, h.size());
    if (rc != SM_OK) return rc;
    SM_HIP(hipMemcpyAsync(s->src.ptr, h.data(), h.size(), hipMemcpyHostToDevice, s->ctx->stream));
    SM_HIP(hipStreamSynchronize(s->ctx->stream));
    return SM_OK;
}

int sm_sh_set_receivers(sm_sh *s, size_t nrec, const size_t *idx, const int *comp, size_t max_steps)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_sh_set_receivers: NULL solver");
    SM_CHECK(nrec == 0 || (idx && comp), SM_ERR_ARG, "sm_sh_set_receivers: NULL array");
    SM_CHECK(nrec < (1u << 24), SM_ERR_ARG, "sm_sh_set_receivers: too many receivers");
    for (size_t r = 0; r < nrec; ++r) {
        SM_CHECK(idx[r] < s->N, SM_ERR_ARG, "receiver %zu: index %zu outside the grid", r, idx[r]);
        SM_CHECK(comp[r] == SM_VY || comp[r] == SM_TXY || comp[r] == SM_TYZ, SM_ERR_ARG,
                 "receiver %zu: component %d not recorded by the SH solver", r, comp[r]);
    }
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    dfree(s->rcv); dfree(s->trace);
    s->nrec = nrec; s->max_steps = max_steps; s->rcv_step0 = s->step;
    if (nrec == 0) return SM_OK;
    std::vector<unsigned char> h(nrec * (sizeof(int64_t) + sizeof(int)));
    for (size_t r = 0; r < nrec; ++r) ((int64_t *)h.data())[r] = (int64_t)idx[r];
    std::memcpy(h.data() + nrec * sizeof(int64_t), comp, nrec * sizeof(int));
    int rc = dalloc(s->rcv, h.size());
    if (rc == SM_OK) rc = dalloc(s->trace, std::max<size_t>(1, nrec * max_steps) * s->elt);
    if (rc != SM_OK) return rc;
    SM_HIP(hipMemcpyAsync(s->rcv.ptr, h.data(), h.size(), hipMemcpyHostToDevice, s->ctx->stream));
    SM_HIP(hipMemsetAsync(s->trace.ptr, 0, s->trace.bytes, s->ctx->stream));
    SM_HIP(hipStreamSynchronize(s->ctx->stream));
    return SM_OK;
}

int sm_sh_run(sm_sh *s, size_t nsteps)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_sh_run: NULL solver");
    SM_CHECK(s->step + nsteps < (1ull << 61), SM_ERR_ARG, "sm_sh_run: step count too large");
    if (nsteps == 0) return SM_OK;
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    return s->elt == 8 ? run<double>(s, nsteps) : run<float>(s, nsteps);
}

int sm_sh_set_timing(sm_sh *s, int on)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_sh_set_timing: NULL solver");
    s->timing = on != 0;
    s->timed_ms = 0.0;
    s->timed_steps = 0;
    return SM_OK;
}

int sm_sh_get_timing(const sm_sh *s, double *ms_total, size_t *steps)
{
    SM_CHECK(s && ms_total && steps, SM_ERR_ARG, "sm_sh_get_timing: NULL argument");
    *ms_total = s->timed_ms;
    *steps = s->timed_steps;
    return SM_OK;
}

int sm_sh_current_step(const sm_sh *s, size_t *it)
{
    SM_CHECK(s && it, SM_ERR_ARG, "sm_sh_current_step: NULL argument");
    *it = s->step;
    return SM_OK;
}

int sm_sh_get_traces(sm_sh *s, size_t it0, size_t count, double *out)
{
    SM_CHECK(s && (count == 0 || out), SM_ERR_ARG, "sm_sh_get_traces: NULL argument");
    SM_CHECK(it0 + count <= s->max_steps, SM_ERR_ARG, "sm_sh_get_traces: rows [%zu, %zu) beyond %zu", it0,
             it0 + count, s->max_steps);
    if (count == 0 || s->nrec == 0) return SM_OK;
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    const char *base = (const char *)s->trace.ptr + it0 * s->nrec * s->elt;
    return s->elt == 8 ? read_back<double>(s, base, count * s->nrec, out)
                       : read_back<float>(s, base, count * s->nrec, out);
}

int sm_sh_get_field(sm_sh *s, int comp, double *out)
{
    SM_CHECK(s && out, SM_ERR_ARG, "sm_sh_get_field: NULL argument");
    const int q = comp == SM_TXY ? W_TXY : comp == SM_TYZ ? W_TYZ : comp == SM_VY ? W_VY : -1;
    SM_CHECK(q >= 0, SM_ERR_ARG, "sm_sh_get_field: component %d is not a field", comp);
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    const char *base = (const char *)s->state.ptr + ((size_t)s->iw * W_NF + q) * s->N * s->elt;
    return s->elt == 8 ? read_back<double>(s, base, s->N, out) : read_back<float>(s, base, s->N, out);
}

int sm_sh_reset(sm_sh *s)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_sh_reset: NULL solver");
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    s->step = 0;
    s->rcv_step0 = 0;
    return zero_state(s);
}

int sm_sh_device_bytes(const sm_sh *s, size_t *bytes)
{
    SM_CHECK(s && bytes, SM_ERR_ARG, "sm_sh_device_bytes: NULL argument");
    size_t b = pstd::grid_device_bytes(s->g);
    for (const sm_dbuf *d : {&s->state, &s->der, &s->mat, &s->mem, &s->src, &s->rcv, &s->trace}) b += d->bytes;
    *bytes = b;
    return SM_OK;
}

}  // extern "C"
