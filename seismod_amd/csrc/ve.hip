// ve.hip -- viscoelastic P-SV waves in a 2D VTI medium (ve_vti_pml):
// pseudo-spectral derivatives on a staggered grid, CPML, fourth-order
// Runge-Kutta in time, L standard-linear-solid mechanisms for the
// quasi-dilatational mode and one for shear.
//
// The reference (ve_vti_pml.c:720-784, propagate.c:1959-2278) evaluates
// four derivative stages per step; each stage takes eight FFT derivatives
// (D_x^+ tau_xx, D_z^- tau_xz, D_x^- tau_xz, D_z^+ tau_zz, D_x^- v_x,
// D_z^- v_z, D_z^+ v_x, D_x^+ v_z), corrects them in the CPML layers and
// forms the pointwise right-hand sides, which the Runge-Kutta combination
// folds into the state.  Here a stage is the eight derivatives (pstd.hpp)
// followed by ONE fused kernel that applies the layer corrections and the
// memory-variable updates, the right-hand sides and the Runge-Kutta update
// of every state value; the step-start copies are replaced by rotating
// four state buffers (w, ws and two for wt, since the relaxation sum reads
// neighbouring cells of the stage input) and by copies folded into the
// layer updates.
//
// The state keeps the reference's layout, (tau_xx, tau_zz, tau_xz, v_x,
// v_z, e1[L], e2, e3) one nx2 x nz2 array after the other, because the
// relaxation sum of tau_xx and tau_zz reads it at flat offset 6 N + n + m
// (propagate.c:2190-2195), e2 rather than e1 when L = 1.  c55_0_ij, the
// factor of e3 in tau_xz (propagate.c:2263), is never set by the reference
// (ve_vti_pml.c:368 allocates it; fresh zero-filled memory), so the term is
// 0 and left out.  Results match the FP64 restatement to FFT rounding.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pstd.hpp"

namespace sm {
namespace {

using pstd::Layer;

enum { W_TXX, W_TZZ, W_TXZ, W_VX, W_VZ, W_E1 };   // then e1[L], e2, e3
enum { M_C11, M_C13, M_C33, M_C55, M_K0, M_C550, M_IRHOI, M_IRHOJ, M_ETA2, M_ETA2IJ, M_TS2, M_TS2IJ, M_NFIX };
// then eta1[L], ts1[L]
enum { L_DX_TXX, L_DZ_TXZ, L_DX_TXZ, L_DZ_TZZ, L_DX_VX, L_DZ_VZ, L_DZ_VX, L_DX_VZ, L_NSET };
constexpr int MAXL = 8;

template <typename T>
struct VeStage {
    const T *d;          // the eight derivatives in L_* order (N each)
    const T *mat;        // M_NFIX + 2L arrays of N
    const T *x;          // stage input state ((7 + L) N)
    const T *ws;         // state at the start of the step
    T *w, *wt;
    T a, b;              // wt = ws + a delta (set_t); w = (acc ? w : ws) + b delta
    int set_t, acc;
    int nx2, nz2, np, L;
    unsigned N;
    Layer<T> Ls[L_NSET];
    pstd::StageSrc<T> src;
};

// add_src_ve_rk4 (src.c:114-142) for source s: Sx to tau_xx, Sz to tau_zz,
// Bulk to both, Sxz to tau_xz, other types nothing
template <typename T>
__device__ __forceinline__ void ve_add_src(T *w, T *wt, size_t N, int s, int nsrc, const int64_t *tab, const T *fct,
                                           T dtw, T dtt, int64_t it)
{
    const int64_t k = 2 * it + 1;
    const int64_t typ = tab[s], cell = tab[nsrc + s], off = tab[2 * nsrc + s], len = tab[3 * nsrc + s];
    if (k >= len) return;
    const T v = fct[off + k];
    auto add = [&](size_t q) {
        if (dtt != T(0)) wt[q + cell] += dtt * v;
        w[q + cell] += dtw * v;
    };
    if (typ == SM_SX || typ == SM_BULK) add(W_TXX * N);
    if (typ == SM_SZ || typ == SM_BULK) add(W_TZZ * N);
    if (typ == SM_SXZ) add(W_TXZ * N);
}

template <typename T>
__device__ __forceinline__ void rk(const VeStage<T> &a, size_t q, T d)
{
    const T s = a.ws[q];
    if (a.set_t) a.wt[q] = s + a.a * d;
    a.w[q] = (a.acc ? a.w[q] : s) + a.b * d;
}

// propagate.c:1959-2278 from the eight derivatives on, then eq. (79)-(83)
template <typename T>
__global__ __launch_bounds__(256) void ve_stage(const VeStage<T> a)
{
    const unsigned n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= a.N) return;
    const size_t N = a.N;
    const int i = (int)(n / (unsigned)a.nz2), j = (int)(n % (unsigned)a.nz2);
    T d[L_NSET];
#pragma unroll
    for (int q = 0; q < L_NSET; ++q) d[q] = pstd::layer_fix(a.Ls[q], a.d[q * N + n], i, j, a.nx2, a.nz2, a.np);
    const T *m = a.mat;
    const int L = a.L;
    // velocities on the interior nodes only (2044-2050, 2107-2112)
    const bool inner = i >= 1 && i < a.nx2 - 1 && j >= 1 && j < a.nz2 - 1;
    rk(a, W_VX * N + n, inner ? m[M_IRHOI * N + n] * (d[L_DX_TXX] + d[L_DZ_TXZ]) : T(0));
    rk(a, W_VZ * N + n, inner ? m[M_IRHOJ * N + n] * (d[L_DX_TXZ] + d[L_DZ_TZZ]) : T(0));
    // relaxation of the dilatation (2172-2180); the sum of tau_xx and
    // tau_zz reads flat offset 6 N + n + k, as written (2190-2195)
    const T t1 = d[L_DX_VX], t2 = d[L_DZ_VZ];
    T se = T(0);
    for (int k = 0; k < L; ++k) se += a.x[6 * N + n + k];
    const T e2 = a.x[(5 + L) * N + n];
    const T c550 = m[M_C550 * N + n], K0 = m[M_K0 * N + n];
    rk(a, W_TXX * N + n, m[M_C11 * N + n] * t1 + m[M_C13 * N + n] * t2 + K0 * se + T(2) * c550 * e2);
    rk(a, W_TZZ * N + n, m[M_C13 * N + n] * t1 + m[M_C33 * N + n] * t2 + K0 * se - T(2) * c550 * e2);
    for (int k = 0; k < L; ++k) {
        const T e1 = a.x[(W_E1 + k) * N + n];
        rk(a, (W_E1 + k) * N + n,
           m[(M_NFIX + L + k) * N + n] * (m[(M_NFIX + k) * N + n] * (t1 + t2) - e1));
    }
    rk(a, (5 + L) * N + n, T(0.5) * m[M_TS2 * N + n] * (m[M_ETA2 * N + n] * (t1 + t2) - T(2) * e2));
    // e3 and tau_xz (2257-2263)
    const T u = d[L_DZ_VX] + d[L_DX_VZ];
    const T e3 = a.x[(6 + L) * N + n];
    rk(a, (6 + L) * N + n, m[M_TS2IJ * N + n] * (m[M_ETA2IJ * N + n] * u - e3));
    rk(a, W_TXZ * N + n, m[M_C55 * N + n] * u);
    for (int s = 0; s < a.src.n; ++s)
        if (a.src.cell[s] == (int)n)
            ve_add_src(a.w, a.wt, N, s, a.src.nsrc, a.src.tab, a.src.fct, a.src.dtw, a.src.dtt, a.src.it);
}

// all sources in order on one thread (more sources than the stage kernel
// takes)
template <typename T>
__global__ void ve_src(T *w, T *wt, size_t N, int nsrc, const int64_t *tab, const T *fct, T dtw, T dtt,
                       int64_t it)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int s = 0; s < nsrc; ++s) ve_add_src(w, wt, N, s, nsrc, tab, fct, dtw, dtt, it);
}

__host__ __device__ inline int comp_slot(int c)
{
    return c == SM_TXX ? W_TXX : c == SM_TZZ ? W_TZZ : c == SM_TXZ ? W_TXZ : c == SM_VX ? W_VX
                                                                        : c == SM_VZ ? W_VZ : -1;
}

template <typename T>
__global__ void ve_gather(const T *w, size_t N, int nrec, const int64_t *ridx, const int *rcomp, T *trace,
                          int64_t row)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrec) return;
    trace[row * nrec + r] = w[(size_t)comp_slot(rcomp[r]) * N + ridx[r]];
}

}  // namespace
}  // namespace sm

using namespace sm;

struct sm_ve {
    sm_ctx *ctx = nullptr;
    sm_ve_params p{};
    pstd::Grid g;
    size_t N = 0, elt = 8, nw = 0;  // nw = 7 + L state arrays
    sm_dbuf state, der, mat;        // 4 x nw N; 8 N; (M_NFIX + 2L) N
    int iw = 0, iws = 1, iwt = 2, iwt2 = 3;
    sm_dbuf mem;                    // 3 copies x (4 dx_* of (2np+1) nz2, 4 dz_* of (2np+1) nx2)
    size_t mem_dx = 0, mem_dz = 0;
    sm_dbuf src, rcv, trace;
    size_t nsrc = 0, nrec = 0, max_steps = 0, rcv_step0 = 0, step = 0;
    size_t src_cells[pstd::SRC_FUSED] = {};   // cells of the sources the stage kernel injects
    bool timing = false;
    double timed_ms = 0.0;
    size_t timed_steps = 0;
};

namespace {

void release(sm_ve *s)
{
    pstd::release(s->g);
    for (sm_dbuf *b : {&s->state, &s->der, &s->mat, &s->mem, &s->src, &s->rcv, &s->trace}) dfree(*b);
}

template <typename T> T *STATE(sm_ve *s, int b) { return (T *)s->state.ptr + (size_t)b * s->nw * s->N; }
// memory set L_*: x sets (DX_*) are (2np+1) nz2, z sets (DZ_*) (2np+1) nx2
bool is_x(int set) { return set == L_DX_TXX || set == L_DX_TXZ || set == L_DX_VX || set == L_DX_VZ; }
template <typename T> T *MEM(sm_ve *s, int copy, int set)
{
    const size_t per = 4 * (s->mem_dx + s->mem_dz);
    size_t off = 0;
    for (int k = 0; k < set; ++k) off += is_x(k) ? s->mem_dx : s->mem_dz;
    return (T *)s->mem.ptr + copy * per + off;
}

// one propagation of x; the fused kernel writes w and the next stage's
// input into S[iwt2] (never x)
template <typename T>
int stage(sm_ve *s, int tabs, bool last, const T *x, T a, T b, int set_t, int acc, T dtw, T dtt, int64_t it)
{
    hipStream_t st = s->ctx->stream;
    const size_t N = s->N;
    const int np = (int)s->g.np;
    T *d = (T *)s->der.ptr;
    // (field, slot, forward) of the derivatives along x, then along z
    const pstd::DJob xj[4] = {{W_TXX, L_DX_TXX, true}, {W_TXZ, L_DX_TXZ, false}, {W_VX, L_DX_VX, false},
                              {W_VZ, L_DX_VZ, true}};
    const pstd::DJob zj[4] = {{W_TXZ, L_DZ_TXZ, false}, {W_TZZ, L_DZ_TZZ, true}, {W_VZ, L_DZ_VZ, false},
                              {W_VX, L_DZ_VX, true}};
    int rc = pstd::deriv_xz<T>(s->g, x, d, xj, 4, zj, 4);
    if (rc != SM_OK) return rc;
    VeStage<T> k;
    k.d = d; k.mat = (const T *)s->mat.ptr; k.x = x; k.ws = STATE<T>(s, s->iws);
    k.w = STATE<T>(s, s->iw); k.wt = STATE<T>(s, s->iwt2);
    k.a = a; k.b = b; k.set_t = set_t; k.acc = acc;
    k.nx2 = s->g.nx2; k.nz2 = s->g.nz2; k.np = np; k.L = s->p.L; k.N = (unsigned)N;
    k.src.n = 0;
    if (s->nsrc && s->nsrc <= (size_t)pstd::SRC_FUSED) {
        k.src.n = (int)s->nsrc;
        for (size_t q = 0; q < s->nsrc; ++q) k.src.cell[q] = (int)s->src_cells[q];
        k.src.tab = (const int64_t *)s->src.ptr;
        k.src.fct = (const T *)(k.src.tab + 4 * s->nsrc);
        k.src.nsrc = (int)s->nsrc;
        k.src.dtw = dtw; k.src.dtt = dtt; k.src.it = it;
    }
    // m2 <- m1 over (2np+1) rows after every propagation (propagate.c:
    // 2265-2277); m0 <- m1 over 2np rows at the start of a step
    // (ve_vti_pml.c:727-738)
    const size_t r2x = (size_t)(2 * np + 1) * s->g.nz2, r2z = (size_t)(2 * np + 1) * s->g.nx2;
    const size_t r0x = last ? (size_t)2 * np * s->g.nz2 : 0, r0z = last ? (size_t)2 * np * s->g.nx2 : 0;
    using namespace pstd;
    auto L3 = [&](int set, int mode, int lastn, int ax, int r1, int k1, int r2, int k2) {
        k.Ls[set] = layer<T>(s->g, tabs, MEM<T>(s, 0, set), MEM<T>(s, 1, set), MEM<T>(s, 2, set), mode, lastn, ax,
                             r1, k1, ax, r2, k2, mode == 0 ? r2x : r2z, mode == 0 ? r0x : r0z);
    };
    // inner half nodes on the first side and outer half nodes (nl <= np)
    // on the far side for the D^+ of stresses at half nodes and of
    // velocities; node tables (nl < np) elsewhere (propagate.c:2000-2254)
    L3(L_DX_TXX, 0, np, 0, T_BH, K_KH, T_BHE, K_KHE);
    L3(L_DZ_TXZ, 1, np - 1, 1, T_B, K_K, T_B, K_K);
    L3(L_DX_TXZ, 0, np - 1, 0, T_B, K_K, T_B, K_K);
    L3(L_DZ_TZZ, 1, np, 1, T_BH, K_KH, T_BHE, K_KHE);
    L3(L_DX_VX, 0, np - 1, 0, T_B, K_K, T_B, K_K);
    L3(L_DZ_VZ, 1, np - 1, 1, T_B, K_K, T_B, K_K);
    L3(L_DZ_VX, 1, np, 1, T_BH, K_KH, T_BHE, K_KHE);
    L3(L_DX_VZ, 0, np, 0, T_BH, K_KH, T_BHE, K_KHE);
    ve_stage<T><<<pstd::blocks(N), 256, 0, st>>>(k);
    return SM_OK;
}

template <typename T>
int run(sm_ve *s, size_t nsteps)
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
        std::swap(s->iw, s->iws);   // Ws = W; stage 1 rewrites all of w
        for (int q = 0; q < 4; ++q) {
            const T *x = q == 0 ? STATE<T>(s, s->iws) : STATE<T>(s, s->iwt);
            int rc = stage<T>(s, stages[q].tabs, q == 3, x, stages[q].a, stages[q].b, q < 3 ? 1 : 0,
                              stages[q].acc, stages[q].sw, stages[q].st, it);
            if (rc != SM_OK) return rc;
            std::swap(s->iwt, s->iwt2);   // the stage's wt is the next input
            if (s->nsrc > (size_t)pstd::SRC_FUSED)
                ve_src<T><<<1, 64, 0, st>>>(STATE<T>(s, s->iw), STATE<T>(s, s->iwt), N, (int)s->nsrc, stab, sfct,
                                            stages[q].sw, stages[q].st, it);
        }
        if (s->timing) SM_HIP(hipEventRecord(ev[2 * k + 1], st));
        if (s->nrec && (size_t)it >= s->rcv_step0 && (size_t)it - s->rcv_step0 < s->max_steps)
            ve_gather<T><<<pstd::blocks(s->nrec), 256, 0, st>>>(STATE<T>(s, s->iw), N, (int)s->nrec, ridx, rcomp,
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

int zero_state(sm_ve *s)
{
    hipStream_t st = s->ctx->stream;
    SM_HIP(hipMemsetAsync(s->state.ptr, 0, s->state.bytes, st));
    SM_HIP(hipMemsetAsync(s->mem.ptr, 0, s->mem.bytes, st));
    if (s->trace.ptr) SM_HIP(hipMemsetAsync(s->trace.ptr, 0, s->trace.bytes, st));
    SM_HIP(hipStreamSynchronize(st));
    s->iw = 0; s->iws = 1; s->iwt = 2; s->iwt2 = 3;
    return SM_OK;
}

template <typename T>
int read_back(sm_ve *s, const void *dev, size_t n, double *out)
{
    std::vector<T> tmp(n);
    SM_HIP(hipMemcpyAsync(tmp.data(), dev, n * sizeof(T), hipMemcpyDeviceToHost, s->ctx->stream));
    SM_HIP(hipStreamSynchronize(s->ctx->stream));
    for (size_t i = 0; i < n; ++i) out[i] = (double)tmp[i];
    return SM_OK;
}

// the material of ve_vti_pml.c:518-683 in M_* order (eta1[L], ts1[L]
// after M_NFIX); the five moduli of write_field_nc into mod when given
bool material(size_t nx2, size_t nz2, int L, const double *Kin, const double *muin, const double *rhoin,
              const double *eps, const double *delta, const double *Q1, const double *f1, const double *Q2,
              const double *f2, std::vector<double> &all, double *const *mod)
{
    const size_t N = nx2 * nz2;
    const double pi = 4.0 * std::atan(1.0);
    std::vector<std::vector<double>> m(M_NFIX + 2 * L, std::vector<double>(N, 0.0));
    std::vector<double> mu(N), rho(N), ts2(N), eta2(N);
    for (size_t n = 0; n < N; ++n) {
        const double K = Kin[n] * 1000.;
        mu[n] = muin[n] * 1000.;
        rho[n] = rhoin[n] * 1.e-6;
        const double c33 = K + 4. / 3. * mu[n];
        const double c55 = mu[n];
        const double c11 = 2. * eps[n] * c33 + c33;
        const double c13 = std::sqrt(2. * delta[n] * c33 * (c33 - c55) + (c33 - c55) * (c33 - c55)) - c55;
        m[M_C11][n] = c11; m[M_C13][n] = c13; m[M_C33][n] = c33;
        double e1s = 0.0;
        for (int k = 0; k < L; ++k) {
            const double q = Q1[(size_t)k * N + n], fr = f1[(size_t)k * N + n];
            const double tau_e = (std::sqrt(q * q + 1) + 1.) / (2. * pi * fr * q);
            const double tau_s = (std::sqrt(q * q + 1) - 1.) / (2. * pi * fr * q);
            const double eta = tau_s / tau_e;
            e1s += eta;
            m[M_NFIX + k][n] = (1. - 1. / eta) / L;
            m[M_NFIX + L + k][n] = 1. / tau_s;
        }
        const double tau_e = (std::sqrt(Q2[n] * Q2[n] + 1) + 1.) / (2. * pi * f2[n] * Q2[n]);
        ts2[n] = (std::sqrt(Q2[n] * Q2[n] + 1) - 1.) / (2. * pi * f2[n] * Q2[n]);
        eta2[n] = ts2[n] / tau_e;
        const double c550 = c55 * eta2[n];
        m[M_C550][n] = c550;
        const double D = 0.5 * (c11 + c33), Kk = D - c55;
        const double c11_0 = c11 - D + Kk * e1s + c55 * eta2[n];
        const double c33_0 = c33 - D + Kk * e1s + c55 * eta2[n];
        m[M_K0][n] = 0.5 * (c11_0 + c33_0) - c550;
    }
    std::vector<double> &ri = m[M_IRHOI], &rj = m[M_IRHOJ], &c55 = m[M_C55], &tij = m[M_TS2IJ], &eij = m[M_ETA2IJ];
    for (size_t i = 0; i + 1 < nx2; ++i)
        for (size_t j = 0; j < nz2; ++j) ri[i * nz2 + j] = 0.5 * (rho[i * nz2 + j] + rho[(i + 1) * nz2 + j]);
    for (size_t i = 0; i < nx2; ++i)
        for (size_t j = 0; j + 1 < nz2; ++j) rj[i * nz2 + j] = 0.5 * (rho[i * nz2 + j] + rho[i * nz2 + j + 1]);
    for (size_t i = 0; i + 1 < nx2; ++i)
        for (size_t j = 0; j + 1 < nz2; ++j) {
            const size_t a = i * nz2 + j, b = (i + 1) * nz2 + j, c = a + 1, d = b + 1;
            c55[a] = 4.0 / (1. / mu[a] + 1. / mu[b] + 1. / mu[c] + 1. / mu[d]);
            tij[a] = 0.25 * (ts2[a] + ts2[b] + ts2[c] + ts2[d]);
            eij[a] = 0.25 * (eta2[a] + eta2[b] + eta2[c] + eta2[d]);
        }
    // last row, then last column (ve_vti_pml.c:610-635)
    for (size_t j = 0; j < nz2; ++j)
        for (std::vector<double> *v : {&ri, &c55, &tij, &eij}) (*v)[(nx2 - 1) * nz2 + j] = (*v)[(nx2 - 2) * nz2 + j];
    for (size_t i = 0; i < nx2; ++i)
        for (std::vector<double> *v : {&rj, &c55, &tij, &eij}) (*v)[i * nz2 + nz2 - 1] = (*v)[i * nz2 + nz2 - 2];
    if (mod) {
        const int src[5] = {M_C11, M_C13, M_C33, M_C55, M_K0};
        for (int q = 0; q < 5; ++q) std::memcpy(mod[q], m[src[q]].data(), N * sizeof(double));
    }
    // stored as 1/rho, 1 - 1/eta, 1/tau_s (650-683)
    for (size_t n = 0; n < N; ++n) {
        ri[n] = 1. / ri[n];
        rj[n] = 1. / rj[n];
        m[M_ETA2][n] = 1. - 1. / eta2[n];
        eij[n] = 1. - 1. / eij[n];
        m[M_TS2][n] = 1. / ts2[n];
        tij[n] = 1. / tij[n];
    }
    all.clear();
    all.reserve(m.size() * N);
    for (auto &a : m) all.insert(all.end(), a.begin(), a.end());
    for (double v : all)
        if (!std::isfinite(v)) return false;
    return true;
}

int check_model(size_t N, int L, const double *K, const double *mu, const double *rho, const double *eps,
                const double *delta, const double *Q1, const double *f1, const double *Q2, const double *f2)
{
    for (size_t n = 0; n < N; ++n) {
        SM_CHECK(rho[n] > 0 && std::isfinite(K[n]) && std::isfinite(mu[n]) && mu[n] > 0 && std::isfinite(eps[n]) &&
                     std::isfinite(delta[n]) && Q2[n] > 0 && f2[n] > 0,
                 SM_ERR_ARG, "cell %zu needs rho, mu, Q2, f2 > 0 and finite K, epsilon, delta", n);
        for (int k = 0; k < L; ++k)
            SM_CHECK(Q1[(size_t)k * N + n] > 0 && f1[(size_t)k * N + n] > 0, SM_ERR_ARG,
                     "cell %zu needs Q1, f1 > 0 (mechanism %d)", n, k);
    }
    return SM_OK;
}

}  // namespace

extern "C" {

size_t sm_ve_bytes_needed(const sm_ve_params *p, size_t nrec, size_t max_steps)
{
    if (!p || (p->storage != SM_STORE_F64 && p->storage != SM_STORE_F32) || p->L < 1 || p->L > MAXL) return 0;
    const size_t e = p->storage == SM_STORE_F64 ? 8 : 4;
    const size_t np = p->abs ? p->np : 0;
    const size_t nx2 = p->nx + 2 * np, nz2 = p->nz + 2 * np, N = nx2 * nz2;
    const size_t nw = 7 + (size_t)p->L;
    const size_t mem = 3 * 4 * (2 * np + 1) * (nx2 + nz2);
    return e * (4 * nw * N + L_NSET * N + (M_NFIX + 2 * (size_t)p->L) * N + mem) +
           pstd::grid_bytes(nx2, nz2, np, e) + e * nrec * max_steps + 12 * nrec;
}

int sm_ve_moduli(size_t nx2, size_t nz2, int L, const double *K, const double *mu, const double *rho,
                 const double *eps, const double *delta, const double *Q1, const double *f1, const double *Q2,
                 const double *f2, double *c11, double *c13, double *c33, double *c55, double *K0)
{
    SM_CHECK(K && mu && rho && eps && delta && Q1 && f1 && Q2 && f2 && c11 && c13 && c33 && c55 && K0, SM_ERR_ARG,
             "sm_ve_moduli: NULL argument");
    SM_CHECK(nx2 >= 2 && nz2 >= 2 && L >= 1 && L <= MAXL, SM_ERR_ARG, "sm_ve_moduli: bad grid or L");
    int rc = check_model(nx2 * nz2, L, K, mu, rho, eps, delta, Q1, f1, Q2, f2);
    if (rc != SM_OK) return rc;
    std::vector<double> all;
    double *const mod[5] = {c11, c13, c33, c55, K0};
    (void)material(nx2, nz2, L, K, mu, rho, eps, delta, Q1, f1, Q2, f2, all, mod);
    return SM_OK;
}

int sm_ve_create(sm_ctx *ctx, const sm_ve_params *pp, const double *K, const double *mu, const double *rho,
                 const double *eps, const double *delta, const double *Q1, const double *f1, const double *Q2,
                 const double *f2, sm_ve **out)
{
    SM_CHECK(ctx && pp && K && mu && rho && eps && delta && Q1 && f1 && Q2 && f2 && out, SM_ERR_ARG,
             "sm_ve_create: NULL argument");
    sm_ve_params p = *pp;
    SM_CHECK(p.storage == SM_STORE_F64 || p.storage == SM_STORE_F32, SM_ERR_ARG,
             "sm_ve_create: storage must be SM_STORE_F64 or SM_STORE_F32");
    SM_CHECK(p.nx >= 2 && p.nz >= 2, SM_ERR_ARG, "sm_ve_create: the model needs at least 2 x 2 nodes");
    SM_CHECK(p.L >= 1 && p.L <= MAXL, SM_ERR_ARG, "sm_ve_create: L must be 1 to %d", MAXL);
    SM_CHECK(p.dx > 0 && p.dz > 0 && p.dt > 0, SM_ERR_ARG, "sm_ve_create: spacings and dt must be > 0");
    SM_CHECK((p.flags & ~SM_PSTD_HIPFFT) == 0, SM_ERR_ARG, "sm_ve_create: unknown flags %#x", p.flags);
    if (!p.abs) p.np = 0;   // io_utils.c:293: no absorbing layers without abs
    const size_t nx2 = p.nx + 2 * p.np, nz2 = p.nz + 2 * p.np, N = nx2 * nz2;
    SM_CHECK(nx2 < (1u << 20) && nz2 < (1u << 20) && N * (7 + (size_t)p.L) < ((size_t)1 << 31), SM_ERR_ARG,
             "sm_ve_create: grid too large");
    int rc = check_model(N, p.L, K, mu, rho, eps, delta, Q1, f1, Q2, f2);
    if (rc != SM_OK) return rc;
    std::vector<double> mat;
    SM_CHECK(material(nx2, nz2, p.L, K, mu, rho, eps, delta, Q1, f1, Q2, f2, mat, nullptr), SM_ERR_ARG,
             "sm_ve_create: material not finite");
    SM_CHECK(set_device(ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");

    sm_ve *s = new sm_ve;
    s->ctx = ctx;
    s->p = p;
    s->N = N;
    s->nw = 7 + (size_t)p.L;
    s->elt = p.storage == SM_STORE_F64 ? 8 : 4;
    const size_t e = s->elt;
    s->mem_dx = (2 * p.np + 1) * nz2;
    s->mem_dz = (2 * p.np + 1) * nx2;
    rc = dalloc(s->state, 4 * s->nw * N * e);
    if (rc == SM_OK) rc = dalloc(s->der, L_NSET * N * e);
    if (rc == SM_OK) rc = dalloc(s->mat, mat.size() * e);
    if (rc == SM_OK) rc = dalloc(s->mem, 3 * 4 * (s->mem_dx + s->mem_dz) * e);
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

int sm_ve_destroy(sm_ve *s)
{
    if (!s) return SM_OK;
    (void)set_device(s->ctx);
    (void)hipStreamSynchronize(s->ctx->stream);
    release(s);
    delete s;
    return SM_OK;
}

int sm_ve_set_sources(sm_ve *s, size_t nsrc, const int *type, const size_t *idx, const size_t *length,
                      const double *fct)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_ve_set_sources: NULL solver");
    SM_CHECK(nsrc == 0 || (type && idx && length && fct), SM_ERR_ARG, "sm_ve_set_sources: NULL array");
    SM_CHECK(nsrc < (1u << 20), SM_ERR_ARG, "sm_ve_set_sources: too many sources");
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

int sm_ve_set_receivers(sm_ve *s, size_t nrec, const size_t *idx, const int *comp, size_t max_steps)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_ve_set_receivers: NULL solver");
    SM_CHECK(nrec == 0 || (idx && comp), SM_ERR_ARG, "sm_ve_set_receivers: NULL array");
    SM_CHECK(nrec < (1u << 24), SM_ERR_ARG, "sm_ve_set_receivers: too many receivers");
    for (size_t r = 0; r < nrec; ++r) {
        SM_CHECK(idx[r] < s->N, SM_ERR_ARG, "receiver %zu: index %zu outside the grid", r, idx[r]);
        SM_CHECK(comp_slot(comp[r]) >= 0, SM_ERR_ARG, "receiver %zu: component %d not recorded by the P-SV solver",
                 r, comp[r]);
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

int sm_ve_run(sm_ve *s, size_t nsteps)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_ve_run: NULL solver");
    SM_CHECK(s->step + nsteps < (1ull << 61), SM_ERR_ARG, "sm_ve_run: step count too large");
    if (nsteps == 0) return SM_OK;
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    return s->elt == 8 ? run<double>(s, nsteps) : run<float>(s, nsteps);
}

int sm_ve_set_timing(sm_ve *s, int on)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_ve_set_timing: NULL solver");
    s->timing = on != 0;
    s->timed_ms = 0.0;
    s->timed_steps = 0;
    return SM_OK;
}

int sm_ve_get_timing(const sm_ve *s, double *ms_total, size_t *steps)
{
    SM_CHECK(s && ms_total && steps, SM_ERR_ARG, "sm_ve_get_timing: NULL argument");
    *ms_total = s->timed_ms;
    *steps = s->timed_steps;
    return SM_OK;
}

int sm_ve_current_step(const sm_ve *s, size_t *it)
{
    SM_CHECK(s && it, SM_ERR_ARG, "sm_ve_current_step: NULL argument");
    *it = s->step;
    return SM_OK;
}

int sm_ve_get_traces(sm_ve *s, size_t it0, size_t count, double *out)
{
    SM_CHECK(s && (count == 0 || out), SM_ERR_ARG, "sm_ve_get_traces: NULL argument");
    SM_CHECK(it0 + count <= s->max_steps, SM_ERR_ARG, "sm_ve_get_traces: rows [%zu, %zu) beyond %zu", it0,
             it0 + count, s->max_steps);
    if (count == 0 || s->nrec == 0) return SM_OK;
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    const char *base = (const char *)s->trace.ptr + it0 * s->nrec * s->elt;
    return s->elt == 8 ? read_back<double>(s, base, count * s->nrec, out)
                       : read_back<float>(s, base, count * s->nrec, out);
}

int sm_ve_get_field(sm_ve *s, int comp, double *out)
{
    SM_CHECK(s && out, SM_ERR_ARG, "sm_ve_get_field: NULL argument");
    const int q = comp_slot(comp);
    SM_CHECK(q >= 0, SM_ERR_ARG, "sm_ve_get_field: component %d is not a field", comp);
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    const char *base = (const char *)s->state.ptr + ((size_t)s->iw * s->nw + q) * s->N * s->elt;
    return s->elt == 8 ? read_back<double>(s, base, s->N, out) : read_back<float>(s, base, s->N, out);
}

int sm_ve_reset(sm_ve *s)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_ve_reset: NULL solver");
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    s->step = 0;
    s->rcv_step0 = 0;
    return zero_state(s);
}

int sm_ve_device_bytes(const sm_ve *s, size_t *bytes)
{
    SM_CHECK(s && bytes, SM_ERR_ARG, "sm_ve_device_bytes: NULL argument");
    size_t b = pstd::grid_device_bytes(s->g);
    for (const sm_dbuf *d : {&s->state, &s->der, &s->mat, &s->mem, &s->src, &s->rcv, &s->trace}) b += d->bytes;
    *bytes = b;
    return SM_OK;
}

}  // extern "C"
