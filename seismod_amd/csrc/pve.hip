// pve.hip -- poroviscoelastic waves in a 2D isotropic medium (pve_iso_pml):
// Biot solid velocity v, fluid filtration velocity q, total stress, fluid
// pressure p and one relaxation variable e (Carcione and Helle 1999),
// pseudo-spectral derivatives on a staggered grid, CPML, fourth-order
// Runge-Kutta, and the stiff viscous coupling of v and q solved
// analytically at the start of every step (eq. 76-77 of Carcione 1996).
//
// The reference (pve_iso_pml.c:1249-1328, propagate.c:884-1329) evaluates
// twelve FFT derivatives per Runge-Kutta stage, corrects them in the CPML
// layers and forms the right-hand sides.  Here a stage is two derivative
// launches (pstd.hpp: all six x derivatives, then all six z derivatives)
// and ONE fused kernel for the layer corrections, memory variables,
// right-hand sides and the Runge-Kutta update; the analytic step is one
// pointwise kernel on the step-start state, and the state buffers rotate
// (w, ws, wt) instead of being copied.  Divergence and curl records take
// spectral derivatives of the last stage's input, the fields the
// reference's FFT plans are bound to (propagate.c:2672-2704).  Results
// match the FP64 restatement to FFT rounding.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pstd.hpp"

namespace sm {
namespace {

using pstd::Layer;

enum { W_TXX, W_TZZ, W_TXZ, W_P, W_VX, W_VZ, W_QX, W_QZ, W_E, W_NF };
// material arrays; the VTI medium stores c11 in M_EM, c13 in M_EMU, alpha1
// in M_ALPHA and adds c33 and alpha3
enum { M_EM, M_EMU, M_M, M_ALPHA, M_TS, M_VPHI, M_IRHOI, M_RFI, M_MI, M_IRHOJ, M_RFJ, M_MJ, M_MUIJ, M_CVI, M_CQI,
       M_CVJ, M_CQJ, M_NMAT, M_C33 = M_NMAT, M_ALPHA3, M_NMAT_VTI };
enum { L_DX_TXX, L_DZ_TXZ, L_DX_P, L_DX_TXZ, L_DZ_TZZ, L_DZ_P, L_DX_VX, L_DZ_VZ, L_DX_QX, L_DZ_QZ, L_DZ_VX,
       L_DX_VZ, L_NSET };
// derivative slots of the divergence and curl records (after the 12 of a stage)
enum { D_DIVX = L_NSET, D_DIVZ, D_CURLX, D_CURLZ, D_NSLOT };

template <typename T>
struct PveStage {
    const T *d;          // the twelve derivatives in L_* order (N each)
    const T *mat;        // M_NMAT (isotropic) or M_NMAT_VTI arrays of N
    const T *x;          // stage input state
    const T *ws;         // state at the start of the step
    T *w, *wt;
    T a, b;              // wt = ws + a delta (set_t); w = (acc ? w : ws) + b delta
    int set_t, acc;
    int nx2, nz2, np;
    unsigned N;
    Layer<T> Ls[L_NSET];
    pstd::StageSrc<T> src;
};

template <typename T>
__device__ __forceinline__ void rk(const PveStage<T> &a, size_t q, T d)
{
    const T s = a.ws[q];
    if (a.set_t) a.wt[q] = s + a.a * d;
    a.w[q] = (a.acc ? a.w[q] : s) + a.b * d;
}

// add_src_rk4 (src.c:77-112) for source s: p (Sf, Bulk), tau_xx (Sx, Bulk,
// Bulk_s), tau_zz (Sz, Bulk, Bulk_s), tau_xz (Sxz)
template <typename T>
__device__ __forceinline__ void pve_add_src(T *w, T *wt, size_t N, int s, int nsrc, const int64_t *tab, const T *fct,
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
    if (typ == SM_SF || typ == SM_BULK) add(W_P * N);
    if (typ == SM_SX || typ == SM_BULK || typ == SM_BULK_S) add(W_TXX * N);
    if (typ == SM_SZ || typ == SM_BULK || typ == SM_BULK_S) add(W_TZZ * N);
    if (typ == SM_SXZ) add(W_TXZ * N);
}

// propagate.c:884-1329 (isotropic) or 1514-1958 (VTI) from the twelve
// derivatives on, then eq. (79)-(83), then the sources on this cell
template <typename T, bool VTI>
__global__ __launch_bounds__(256) void pve_stage(const PveStage<T> a)
{
    const unsigned n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= a.N) return;
    const size_t N = a.N;
    const int i = (int)(n / (unsigned)a.nz2), j = (int)(n % (unsigned)a.nz2);
    T d[L_NSET];
#pragma unroll
    for (int q = 0; q < L_NSET; ++q) d[q] = pstd::layer_fix(a.Ls[q], a.d[q * N + n], i, j, a.nx2, a.nz2, a.np);
    const T *m = a.mat;
    // solid and fluid velocities on the interior nodes only (998-1005,
    // 1091-1098); elsewhere their rates stay 0
    const bool inner = i >= 1 && i < a.nx2 - 1 && j >= 1 && j < a.nz2 - 1;
    T dvx = T(0), dqx = T(0), dvz = T(0), dqz = T(0);
    if (inner) {
        const T t1 = d[L_DX_TXX] + d[L_DZ_TXZ];
        dqx = (d[L_DX_P] + m[M_RFI * N + n] * t1) / m[M_MI * N + n];
        dvx = m[M_IRHOI * N + n] * t1 - m[M_RFI * N + n] * dqx;
        const T t2 = d[L_DX_TXZ] + d[L_DZ_TZZ];
        dqz = (d[L_DZ_P] + m[M_RFJ * N + n] * t2) / m[M_MJ * N + n];
        dvz = m[M_IRHOJ * N + n] * t2 - m[M_RFJ * N + n] * dqz;
    }
    rk(a, W_VX * N + n, dvx);
    rk(a, W_QX * N + n, dqx);
    rk(a, W_VZ * N + n, dvz);
    rk(a, W_QZ * N + n, dqz);
    // epsilon = alpha (D_x^- v_x + D_z^- v_z) + D_x^- q_x + D_z^- q_z
    // (1100-1211); e, p, tau_xx, tau_zz (1213-1251)
    const T alpha = m[M_ALPHA * N + n];
    T eps;
    if (VTI) eps = alpha * d[L_DX_VX] + m[M_ALPHA3 * N + n] * d[L_DZ_VZ];
    else eps = alpha * (d[L_DX_VX] + d[L_DZ_VZ]);
    eps += d[L_DX_QX];
    eps += d[L_DZ_QZ];
    const T e = a.x[W_E * N + n];
    rk(a, W_E * N + n, -m[M_TS * N + n] * (m[M_VPHI * N + n] * eps + e));
    const T eps2 = m[M_M * N + n] * eps + e;
    rk(a, W_P * N + n, -eps2);
    const T E = m[M_EM * N + n], Emu = m[M_EMU * N + n];
    if (VTI) {   // c11, c13, c33, alpha1, alpha3 (propagate.c:1864-1880)
        rk(a, W_TXX * N + n, E * d[L_DX_VX] + Emu * d[L_DZ_VZ] + alpha * eps2);
        rk(a, W_TZZ * N + n, Emu * d[L_DX_VX] + m[M_C33 * N + n] * d[L_DZ_VZ] + m[M_ALPHA3 * N + n] * eps2);
    } else {
        rk(a, W_TXX * N + n, E * d[L_DX_VX] + Emu * d[L_DZ_VZ] + alpha * eps2);
        rk(a, W_TZZ * N + n, Emu * d[L_DX_VX] + E * d[L_DZ_VZ] + alpha * eps2);
    }
    // tau_xz (1259-1311)
    rk(a, W_TXZ * N + n, m[M_MUIJ * N + n] * (d[L_DZ_VX] + d[L_DX_VZ]));
    for (int s = 0; s < a.src.n; ++s)
        if (a.src.cell[s] == (int)n)
            pve_add_src(a.w, a.wt, N, s, a.src.nsrc, a.src.tab, a.src.fct, a.src.dtw, a.src.dtt, a.src.it);
}

// the analytic step of pve_iso_pml.c:1255-1262 on the step-start state:
// v <- v + coeff_v q, q <- coeff_q q
template <typename T>
__global__ void pve_split(T *ws, const T *mat, unsigned N)
{
    const unsigned n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const size_t NN = N;
    const T qx = ws[W_QX * NN + n], qz = ws[W_QZ * NN + n];
    ws[W_VX * NN + n] = ws[W_VX * NN + n] + mat[M_CVI * NN + n] * qx;
    ws[W_VZ * NN + n] = ws[W_VZ * NN + n] + mat[M_CVJ * NN + n] * qz;
    ws[W_QX * NN + n] = mat[M_CQI * NN + n] * qx;
    ws[W_QZ * NN + n] = mat[M_CQJ * NN + n] * qz;
}

// all sources in order on one thread (more sources than the stage kernel
// takes)
template <typename T>
__global__ void pve_src(T *w, T *wt, size_t N, int nsrc, const int64_t *tab, const T *fct, T dtw, T dtt,
                        int64_t it)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int s = 0; s < nsrc; ++s) pve_add_src(w, wt, N, s, nsrc, tab, fct, dtw, dtt, it);
}

__host__ __device__ inline int comp_slot(int c)
{
    switch (c) {
    case SM_TXX: return W_TXX;
    case SM_TZZ: return W_TZZ;
    case SM_TXZ: return W_TXZ;
    case SM_P: return W_P;
    case SM_VX: return W_VX;
    case SM_VZ: return W_VZ;
    case SM_QX: return W_QX;
    case SM_QZ: return W_QZ;
    default: return -1;
    }
}

// traces: state components from w, divergence and curl from their
// derivative slots (dd = the derivative fields)
template <typename T>
__global__ void pve_gather(const T *w, const T *dd, size_t N, int nrec, const int64_t *ridx, const int *rcomp,
                           T *trace, int64_t row)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrec) return;
    const int c = rcomp[r];
    const size_t n = (size_t)ridx[r];
    T v;
    if (c == SM_DIV) v = dd[D_DIVX * N + n] + dd[D_DIVZ * N + n];
    else if (c == SM_CURL) v = dd[D_CURLX * N + n] - dd[D_CURLZ * N + n];
    else v = w[(size_t)comp_slot(c) * N + n];
    trace[row * nrec + r] = v;
}

// Kinetic energy of a region, summed in FP64 in a fixed order (one partial
// sum per block, then one block over the partials): per cell 0.5 a_x v_x^2
// + b_x v_x q_x + 0.5 c_x q_x^2 and the same along z, coef = a_x, b_x, c_x,
// a_z, b_z, c_z of the region's ni x nj cells (i outer) one array after the
// other.
constexpr unsigned E_BLOCKS = 256;

__device__ inline double block_sum(double a, double *red)
{
    red[threadIdx.x] = a;
    __syncthreads();
    for (unsigned k = blockDim.x / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    return red[0];
}

template <typename T>
__global__ __launch_bounds__(256) void pve_energy_part(const T *w, size_t N, const double *coef, unsigned ni,
                                                      unsigned nj, unsigned i1, unsigned j1, unsigned nz2,
                                                      double *part)
{
    __shared__ double red[256];
    const unsigned ncell = ni * nj;
    double acc = 0.0;
    for (unsigned c = blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += gridDim.x * blockDim.x) {
        const size_t n = (size_t)(c / nj + i1) * nz2 + c % nj + j1;
        const double vx = w[W_VX * N + n], vz = w[W_VZ * N + n], qx = w[W_QX * N + n], qz = w[W_QZ * N + n];
        const double *k = coef + c;
        const size_t nc = ncell;
        acc += 0.5 * k[0] * vx * vx + k[nc] * vx * qx + 0.5 * k[2 * nc] * qx * qx + 0.5 * k[3 * nc] * vz * vz +
               k[4 * nc] * vz * qz + 0.5 * k[5 * nc] * qz * qz;
    }
    const double t = block_sum(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void pve_energy_sum(const double *part, unsigned nb, double *out)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (unsigned b = threadIdx.x; b < nb; b += blockDim.x) acc += part[b];
    const double t = block_sum(acc, red);
    if (threadIdx.x == 0) out[0] = t;
}

}  // namespace
}  // namespace sm

using namespace sm;

struct sm_pve {
    sm_ctx *ctx = nullptr;
    sm_pve_params p{};
    pstd::Grid g;
    size_t N = 0, elt = 8;
    sm_dbuf state, der, mat;        // 3 x 9 N; D_NSLOT N; M_NMAT (M_NMAT_VTI) N
    bool vti = false;
    int iw = 0, iws = 1, iwt = 2;
    sm_dbuf mem;                    // 3 copies x (6 dx_* of (2np+1) nz2, 6 dz_* of (2np+1) nx2)
    size_t mem_dx = 0, mem_dz = 0;
    sm_dbuf src, rcv, trace;
    size_t nsrc = 0, nrec = 0, max_steps = 0, rcv_step0 = 0, step = 0;
    size_t src_cells[pstd::SRC_FUSED] = {};   // cells of the sources the stage kernel injects
    // kinetic energy of a region: coefficients, partial sums, one value per step
    sm_dbuf ecoef, epart, etrace;
    size_t ei1 = 0, ej1 = 0, eni = 0, enj = 0, e_max = 0, e_step0 = 0;
    bool need_div = false, need_curl = false;
    bool timing = false;
    double timed_ms = 0.0;
    size_t timed_steps = 0;
};

namespace {

void release(sm_pve *s)
{
    pstd::release(s->g);
    for (sm_dbuf *b : {&s->state, &s->der, &s->mat, &s->mem, &s->src, &s->rcv, &s->trace, &s->ecoef, &s->epart,
                       &s->etrace})
        dfree(*b);
}

template <typename T> T *STATE(sm_pve *s, int b) { return (T *)s->state.ptr + (size_t)b * W_NF * s->N; }
bool is_x(int set)
{
    return set == L_DX_TXX || set == L_DX_P || set == L_DX_TXZ || set == L_DX_VX || set == L_DX_QX ||
           set == L_DX_VZ;
}
template <typename T> T *MEM(sm_pve *s, int copy, int set)
{
    const size_t per = 6 * (s->mem_dx + s->mem_dz);
    size_t off = 0;
    for (int k = 0; k < set; ++k) off += is_x(k) ? s->mem_dx : s->mem_dz;
    return (T *)s->mem.ptr + copy * per + off;
}

// divergence and curl derivatives of x (the last stage's input) into
// their slots
template <typename T>
int div_curl(sm_pve *s, const T *x, bool div, bool curl)
{
    T *d = (T *)s->der.ptr;
    pstd::DJob xj[2], zj[2];
    int nx = 0, nz = 0;
    if (div) { xj[nx++] = {W_VX, D_DIVX, false}; zj[nz++] = {W_VZ, D_DIVZ, false}; }
    if (curl) { xj[nx++] = {W_VZ, D_CURLX, true}; zj[nz++] = {W_VX, D_CURLZ, true}; }
    return pstd::deriv_xz<T>(s->g, x, d, xj, nx, zj, nz);
}

template <typename T>
int stage(sm_pve *s, int tabs, bool last, const T *x, T a, T b, int set_t, int acc, T dtw, T dtt, int64_t it)
{
    hipStream_t st = s->ctx->stream;
    const size_t N = s->N;
    const int np = (int)s->g.np;
    T *d = (T *)s->der.ptr;
    // (field, slot, forward) along x, then along z (propagate.c:909-1300)
    const pstd::DJob xj[6] = {{W_TXX, L_DX_TXX, true}, {W_P, L_DX_P, true},   {W_TXZ, L_DX_TXZ, false},
                              {W_VX, L_DX_VX, false},  {W_QX, L_DX_QX, false}, {W_VZ, L_DX_VZ, true}};
    const pstd::DJob zj[6] = {{W_TXZ, L_DZ_TXZ, false}, {W_TZZ, L_DZ_TZZ, true}, {W_P, L_DZ_P, true},
                              {W_VZ, L_DZ_VZ, false},   {W_QZ, L_DZ_QZ, false},  {W_VX, L_DZ_VX, true}};
    int rc = pstd::deriv_xz<T>(s->g, x, d, xj, 6, zj, 6);
    if (rc != SM_OK) return rc;
    PveStage<T> k;
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
    // m2 <- m1 over (2np+1) rows after every propagation (propagate.c:
    // 1313-1328 and 1940-1957); m0 <- m1 at the start of a step, over 2np+1
    // rows (pve_iso_pml.c:1270-1285) or 2np rows (pve_vti_pml.c:1226-1241:
    // the far layers' innermost row of m0 then stays 0), folded into the
    // last stage
    const size_t r2x = (size_t)(2 * np + 1) * s->g.nz2, r2z = (size_t)(2 * np + 1) * s->g.nx2;
    const size_t r0rows = s->vti ? (size_t)(2 * np) : (size_t)(2 * np + 1);
    const size_t r0x = last ? r0rows * s->g.nz2 : 0, r0z = last ? r0rows * s->g.nx2 : 0;
    using namespace pstd;
    auto L3 = [&](int set, int mode, int lastn, int ax, int r1, int k1, int r2, int k2) {
        k.Ls[set] = layer<T>(s->g, tabs, MEM<T>(s, 0, set), MEM<T>(s, 1, set), MEM<T>(s, 2, set), mode, lastn, ax,
                             r1, k1, ax, r2, k2, mode == 0 ? r2x : r2z, mode == 0 ? r0x : r0z);
    };
    L3(L_DX_TXX, 0, np, 0, T_BH, K_KH, T_BHE, K_KHE);
    L3(L_DZ_TXZ, 1, np - 1, 1, T_B, K_K, T_B, K_K);
    L3(L_DX_P, 0, np, 0, T_BH, K_KH, T_BHE, K_KHE);
    L3(L_DX_TXZ, 0, np - 1, 0, T_B, K_K, T_B, K_K);
    L3(L_DZ_TZZ, 1, np, 1, T_BH, K_KH, T_BHE, K_KHE);
    L3(L_DZ_P, 1, np, 1, T_BH, K_KH, T_BHE, K_KHE);
    L3(L_DX_VX, 0, np - 1, 0, T_B, K_K, T_B, K_K);
    L3(L_DZ_VZ, 1, np - 1, 1, T_B, K_K, T_B, K_K);
    L3(L_DX_QX, 0, np - 1, 0, T_B, K_K, T_B, K_K);
    L3(L_DZ_QZ, 1, np - 1, 1, T_B, K_K, T_B, K_K);
    L3(L_DZ_VX, 1, np, 1, T_BH, K_KH, T_BHE, K_KHE);
    L3(L_DX_VZ, 0, np, 0, T_BH, K_KH, T_BHE, K_KHE);
    if (s->vti) pve_stage<T, true><<<pstd::blocks(N), 256, 0, st>>>(k);
    else pve_stage<T, false><<<pstd::blocks(N), 256, 0, st>>>(k);
    return SM_OK;
}

template <typename T>
int run(sm_pve *s, size_t nsteps)
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
    const T h6 = dt / T(6), h3 = dt / T(3), h2 = dt / T(2);
    struct Stage { int tabs; T a; T b; int acc; T sw; T st; };
    const Stage stages[4] = {{0, h2, h6, 0, h6, h2}, {1, h2, h3, 1, h3, h2},
                             {1, dt, h3, 1, h3, dt}, {2, T(0), h6, 1, h6, T(0)}};
    for (size_t k = 0; k < nsteps; ++k) {
        const int64_t it = (int64_t)(s->step + k);
        if (s->timing) SM_HIP(hipEventRecord(ev[2 * k], st));
        std::swap(s->iw, s->iws);   // Ws = W, then the analytic step on it
        pve_split<T><<<pstd::blocks(N), 256, 0, st>>>(STATE<T>(s, s->iws), (const T *)s->mat.ptr, (unsigned)N);
        for (int q = 0; q < 4; ++q) {
            const T *x = q == 0 ? STATE<T>(s, s->iws) : STATE<T>(s, s->iwt);
            int rc = stage<T>(s, stages[q].tabs, q == 3, x, stages[q].a, stages[q].b, q < 3 ? 1 : 0,
                              stages[q].acc, stages[q].sw, stages[q].st, it);
            if (rc != SM_OK) return rc;
            if (s->nsrc > (size_t)pstd::SRC_FUSED)
                pve_src<T><<<1, 64, 0, st>>>(STATE<T>(s, s->iw), STATE<T>(s, s->iwt), N, (int)s->nsrc, stab, sfct,
                                             stages[q].sw, stages[q].st, it);
        }
        if (s->timing) SM_HIP(hipEventRecord(ev[2 * k + 1], st));
        if (s->nrec && (size_t)it >= s->rcv_step0 && (size_t)it - s->rcv_step0 < s->max_steps) {
            if (s->need_div || s->need_curl) {
                int rc = div_curl<T>(s, STATE<T>(s, s->iwt), s->need_div, s->need_curl);
                if (rc != SM_OK) return rc;
            }
            pve_gather<T><<<pstd::blocks(s->nrec), 256, 0, st>>>(STATE<T>(s, s->iw), (const T *)s->der.ptr, N,
                                                                 (int)s->nrec, ridx, rcomp, (T *)s->trace.ptr,
                                                                 (int64_t)((size_t)it - s->rcv_step0));
        }
        if (s->e_max && (size_t)it >= s->e_step0 && (size_t)it - s->e_step0 < s->e_max) {
            const size_t nc = s->eni * s->enj;
            const unsigned nb = (unsigned)std::min<size_t>(E_BLOCKS, (nc + 255) / 256);
            pve_energy_part<T><<<nb, 256, 0, st>>>(STATE<T>(s, s->iw), N, (const double *)s->ecoef.ptr,
                                                   (unsigned)s->eni, (unsigned)s->enj, (unsigned)s->ei1,
                                                   (unsigned)s->ej1, (unsigned)s->g.nz2, (double *)s->epart.ptr);
            pve_energy_sum<<<1, 256, 0, st>>>((const double *)s->epart.ptr, nb,
                                              (double *)s->etrace.ptr + ((size_t)it - s->e_step0));
        }
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

int zero_state(sm_pve *s)
{
    hipStream_t st = s->ctx->stream;
    SM_HIP(hipMemsetAsync(s->state.ptr, 0, s->state.bytes, st));
    SM_HIP(hipMemsetAsync(s->mem.ptr, 0, s->mem.bytes, st));
    SM_HIP(hipMemsetAsync(s->der.ptr, 0, s->der.bytes, st));
    if (s->trace.ptr) SM_HIP(hipMemsetAsync(s->trace.ptr, 0, s->trace.bytes, st));
    if (s->etrace.ptr) SM_HIP(hipMemsetAsync(s->etrace.ptr, 0, s->etrace.bytes, st));
    SM_HIP(hipStreamSynchronize(st));
    s->iw = 0; s->iws = 1; s->iwt = 2;
    return SM_OK;
}

template <typename T>
int read_back(sm_pve *s, const void *dev, size_t n, double *out)
{
    std::vector<T> tmp(n);
    SM_HIP(hipMemcpyAsync(tmp.data(), dev, n * sizeof(T), hipMemcpyDeviceToHost, s->ctx->stream));
    SM_HIP(hipStreamSynchronize(s->ctx->stream));
    for (size_t i = 0; i < n; ++i) out[i] = (double)tmp[i];
    return SM_OK;
}

// model inputs of either medium (the VTI ones null for the isotropic one)
struct PveModel {
    const double *K_m, *K_s, *K_f, *phi, *mu, *rho_s, *rho_f, *T1, *T3, *eta, *kappa1, *kappa3, *Q, *fq, *eps,
        *delta;
};

// pml_median_average (pml.c:2254-2329): ring by ring from the layer next
// to the model outwards, each ring node replaced by the lower median of its
// 3 x 3 neighbourhood inside the grid, from the previous ring's values (the
// window stays 3 x 3: the size update at pml.c:2324 assigns nmed its own old
// value); a ring's columns stop at nz2 - np - n - 1 (pml.c:2297)
void pml_median(std::vector<double> &a, size_t nx2, size_t nz2, size_t np)
{
    std::vector<double> out;
    double w[9];
    auto med = [&](long i, long j) {
        int k = 0;
        for (long ii = i - 1; ii <= i + 1; ++ii)
            for (long jj = j - 1; jj <= j + 1; ++jj)
                if (ii >= 0 && ii < (long)nx2 && jj >= 0 && jj < (long)nz2) w[k++] = a[(size_t)ii * nz2 + jj];
        std::nth_element(w, w + (k - 1) / 2, w + k);
        return w[(k - 1) / 2];
    };
    for (long n = 1; n <= (long)np; ++n) {
        const long d = (long)np - n;
        out = a;
        for (long i = d; i < (long)nx2 - d; ++i) {
            out[(size_t)i * nz2 + d] = med(i, d);
            out[(size_t)i * nz2 + nz2 - d - 1] = med(i, (long)nz2 - d - 1);
        }
        for (long j = d + 1; j < (long)nz2 - (long)np - n - 1; ++j) {
            out[(size_t)d * nz2 + j] = med(d, j);
            out[(size_t)(nx2 - d - 1) * nz2 + j] = med((long)nx2 - d - 1, j);
        }
        a.swap(out);
    }
}

// the material of pve_iso_pml.c:795-1145 or pve_vti_pml.c:787-1145 in M_*
// order; median: the smoothing of the isotropic medium's layers (833-845)
bool material(size_t nx2, size_t nz2, double dt, bool vti, const PveModel &in, size_t median_np,
              std::vector<double> &all)
{
    const size_t N = nx2 * nz2;
    const double pi = 4.0 * std::atan(1.0);
    std::vector<std::vector<double>> m(vti ? M_NMAT_VTI : M_NMAT, std::vector<double>(N, 0.0));
    std::vector<double> mu(N), rho_f(N), rho(N), m1(N), m3(N), nk1(N), nk3(N), tau_s(N), vphi(N), eta(N), k1(N),
        k3(N);
    for (size_t n = 0; n < N; ++n) {
        const double K_m = in.K_m[n] * 1000., K_s = in.K_s[n] * 1000., K_f = in.K_f[n] * 1000.;
        mu[n] = in.mu[n] * 1000.;
        eta[n] = in.eta[n] * 1.013249965828145e6;
        k1[n] = in.kappa1[n];
        k3[n] = in.kappa3[n];
        const double rho_s = in.rho_s[n] * 1.e-6;
        rho_f[n] = in.rho_f[n] * 1.e-6;
        const double phi = in.phi[n];
        const double tau0 = 1. / (2. * pi * in.fq[n]);
        const double D = K_s * (1. + phi * (K_s / K_f - 1.));
        if (vti) {   // eq. 9-15 of Carcione 1996, Thomsen's c11 and c13
            const double c33 = K_m + 4. / 3. * mu[n], c55 = mu[n];
            const double c11 = 2. * in.eps[n] * c33 + c33, c12 = c11 - 2. * c55;
            const double c13 = std::sqrt(2. * in.delta[n] * c33 * (c33 - c55) + (c33 - c55) * (c33 - c55)) - c55;
            m[M_EM][n] = c11;
            m[M_EMU][n] = c13;
            m[M_C33][n] = c33;
            m[M_M][n] = K_s * K_s / (D - (2. * c11 + c33 + 2. * c12 + 4. * c13) / 9.);
            m[M_ALPHA][n] = 1. - 1. / 3. * (c11 + c12 + c13) / K_s;
            m[M_ALPHA3][n] = 1. - 1. / 3. * (2. * c13 + c33) / K_s;
        } else {
            m[M_EM][n] = K_m + 4. / 3. * mu[n];
            m[M_M][n] = K_s * K_s / (D - K_m);
            m[M_ALPHA][n] = 1. - K_m / K_s;
        }
        m1[n] = in.T1[n] * rho_f[n] / phi;
        m3[n] = in.T3[n] * rho_f[n] / phi;
        const double Q = in.Q[n];
        if (Q < 1.e4) {
            const double tau_e = tau0 / Q * (std::sqrt(Q * Q + 1.) + 1.);
            tau_s[n] = tau0 / Q * (std::sqrt(Q * Q + 1.) - 1.);
            vphi[n] = tau_e / tau_s[n] - 1.;
        } else {
            tau_s[n] = tau0;
            vphi[n] = 0.0;
        }
        rho[n] = (1. - phi) * rho_s + phi * rho_f[n];
    }
    if (median_np && !vti) {
        for (std::vector<double> *v : {&mu, &m[M_EM], &m[M_M], &m[M_ALPHA], &vphi, &tau_s, &rho, &rho_f, &eta, &k1, &m1})
            pml_median(*v, nx2, nz2, median_np);
        k3 = k1;
        m3 = m1;
    }
    for (size_t n = 0; n < N; ++n) {
        nk1[n] = eta[n] / k1[n];
        nk3[n] = eta[n] / k3[n];
    }
    auto coeffs = [&](double r, double rf, double mi, double nk, double &cq, double &cv) {
        const double tmp = 1. / (rf * rf - r * mi);
        const double beta12 = tmp * rf, beta22 = tmp * -r;
        const double lam = -nk * beta22;
        cq = std::exp(lam * dt);
        cv = beta12 / beta22 * (cq - 1.);
    };
    std::vector<double> rho_i(N), rho_f_i(N), m_i(N), rho_j(N), rho_f_j(N), m_j(N);
    for (size_t i = 0; i + 1 < nx2; ++i)
        for (size_t j = 0; j < nz2; ++j) {
            const size_t a = i * nz2 + j, b = (i + 1) * nz2 + j;
            rho_i[a] = 0.5 * (rho[a] + rho[b]);
            rho_f_i[a] = 0.5 * (rho_f[a] + rho_f[b]);
            const double nk = 0.5 * (nk1[a] + nk1[b]);
            m_i[a] = 0.5 * (m1[a] + m1[b]);
            coeffs(rho_i[a], rho_f_i[a], m_i[a], nk, m[M_CQI][a], m[M_CVI][a]);
        }
    for (size_t i = 0; i < nx2; ++i)
        for (size_t j = 0; j + 1 < nz2; ++j) {
            const size_t a = i * nz2 + j, b = a + 1;
            rho_j[a] = 0.5 * (rho[a] + rho[b]);
            rho_f_j[a] = 0.5 * (rho_f[a] + rho_f[b]);
            const double nk = 0.5 * (nk3[a] + nk3[b]);
            m_j[a] = 0.5 * (m3[a] + m3[b]);
            coeffs(rho_j[a], rho_f_j[a], m_j[a], nk, m[M_CQJ][a], m[M_CVJ][a]);
        }
    std::vector<double> &muij = m[M_MUIJ];
    for (size_t i = 0; i + 1 < nx2; ++i)
        for (size_t j = 0; j + 1 < nz2; ++j) {
            const size_t a = i * nz2 + j, b = (i + 1) * nz2 + j, c = a + 1, d = b + 1;
            muij[a] = 4.0 / (1. / mu[a] + 1. / mu[b] + 1. / mu[c] + 1. / mu[d]);
        }
    // last row, then last column (pve_iso_pml.c:1076-1105)
    for (size_t j = 0; j < nz2; ++j)
        for (std::vector<double> *v : {&rho_i, &rho_f_i, &m_i, &m[M_CQI], &m[M_CVI], &muij})
            (*v)[(nx2 - 1) * nz2 + j] = (*v)[(nx2 - 2) * nz2 + j];
    for (size_t i = 0; i < nx2; ++i)
        for (std::vector<double> *v : {&rho_j, &rho_f_j, &m_j, &m[M_CQJ], &m[M_CVJ], &muij})
            (*v)[i * nz2 + nz2 - 1] = (*v)[i * nz2 + nz2 - 2];
    // stored forms (pve_iso_pml.c:1131-1145)
    for (size_t n = 0; n < N; ++n) {
        m[M_TS][n] = 1. / tau_s[n];
        m[M_VPHI][n] = m[M_M][n] * vphi[n] / (1. + vphi[n]);
        if (!vti) m[M_EMU][n] = m[M_EM][n] - 2. * mu[n];
        m[M_MI][n] = rho_f_i[n] * rho_f_i[n] / rho_i[n] - m_i[n];
        m[M_MJ][n] = rho_f_j[n] * rho_f_j[n] / rho_j[n] - m_j[n];
        m[M_RFI][n] = rho_f_i[n] / rho_i[n];
        m[M_RFJ][n] = rho_f_j[n] / rho_j[n];
        m[M_IRHOI][n] = 1. / rho_i[n];
        m[M_IRHOJ][n] = 1. / rho_j[n];
    }
    all.clear();
    all.reserve(m.size() * N);
    for (auto &a : m) all.insert(all.end(), a.begin(), a.end());
    for (double v : all)
        if (!std::isfinite(v)) return false;
    return true;
}

int create(sm_ctx *ctx, const sm_pve_params *pp, bool vti, const PveModel &in, sm_pve **out)
{
    const char *fn = vti ? "sm_pve_create_vti" : "sm_pve_create";
    SM_CHECK(ctx && pp && out, SM_ERR_ARG, "%s: NULL argument", fn);
    SM_CHECK(in.K_m && in.K_s && in.K_f && in.phi && in.mu && in.rho_s && in.rho_f && in.T1 && in.T3 && in.eta &&
                 in.kappa1 && in.kappa3 && in.Q && in.fq && (!vti || (in.eps && in.delta)),
             SM_ERR_ARG, "%s: NULL model array", fn);
    sm_pve_params p = *pp;
    SM_CHECK(p.storage == SM_STORE_F64 || p.storage == SM_STORE_F32, SM_ERR_ARG,
             "%s: storage must be SM_STORE_F64 or SM_STORE_F32", fn);
    SM_CHECK(p.nx >= 2 && p.nz >= 2, SM_ERR_ARG, "%s: the model needs at least 2 x 2 nodes", fn);
    SM_CHECK(p.dx > 0 && p.dz > 0 && p.dt > 0, SM_ERR_ARG, "%s: spacings and dt must be > 0", fn);
    SM_CHECK((p.flags & ~(SM_PSTD_HIPFFT | SM_PVE_PML_MEDIAN)) == 0, SM_ERR_ARG, "%s: unknown flags %#x", fn, p.flags);
    SM_CHECK(!(vti && (p.flags & SM_PVE_PML_MEDIAN)), SM_ERR_ARG,
             "%s: the median smoothing of the layers is for the isotropic medium", fn);
    if (!p.abs) p.np = 0;   // io_utils.c:293: no absorbing layers without abs
    const size_t nx2 = p.nx + 2 * p.np, nz2 = p.nz + 2 * p.np, N = nx2 * nz2;
    SM_CHECK(nx2 < (1u << 20) && nz2 < (1u << 20) && N * W_NF < ((size_t)1 << 31), SM_ERR_ARG, "%s: grid too large",
             fn);
    for (size_t n = 0; n < N; ++n)
        SM_CHECK(in.K_m[n] >= 0 && in.K_s[n] > 0 && in.K_f[n] > 0 && in.phi[n] > 0 && in.mu[n] >= 0 &&
                     in.rho_s[n] > 0 && in.rho_f[n] > 0 && in.T1[n] > 0 && in.T3[n] > 0 && in.eta[n] >= 0 &&
                     in.kappa1[n] > 0 && in.kappa3[n] > 0 && in.Q[n] > 0 && in.fq[n] > 0,
                 SM_ERR_ARG, "%s: cell %zu has a property out of range", fn, n);
    std::vector<double> mat;
    SM_CHECK(material(nx2, nz2, p.dt, vti, in, (p.flags & SM_PVE_PML_MEDIAN) ? p.np : 0, mat), SM_ERR_ARG,
             "%s: material not finite", fn);
    SM_CHECK(set_device(ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");

    sm_pve *s = new sm_pve;
    s->ctx = ctx;
    s->p = p;
    s->N = N;
    s->vti = vti;
    s->elt = p.storage == SM_STORE_F64 ? 8 : 4;
    const size_t e = s->elt;
    s->mem_dx = (2 * p.np + 1) * nz2;
    s->mem_dz = (2 * p.np + 1) * nx2;
    int rc = dalloc(s->state, 3 * W_NF * N * e);
    if (rc == SM_OK) rc = dalloc(s->der, D_NSLOT * N * e);
    if (rc == SM_OK) rc = dalloc(s->mat, mat.size() * e);
    if (rc == SM_OK) rc = dalloc(s->mem, 3 * 6 * (s->mem_dx + s->mem_dz) * e);
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

}  // namespace

extern "C" {

size_t sm_pve_bytes_needed(const sm_pve_params *p, size_t nrec, size_t max_steps)
{
    if (!p || (p->storage != SM_STORE_F64 && p->storage != SM_STORE_F32)) return 0;
    const size_t e = p->storage == SM_STORE_F64 ? 8 : 4;
    const size_t np = p->abs ? p->np : 0;
    const size_t nx2 = p->nx + 2 * np, nz2 = p->nz + 2 * np, N = nx2 * nz2;
    const size_t mem = 3 * 6 * (2 * np + 1) * (nx2 + nz2);
    return e * (3 * W_NF * N + D_NSLOT * N + M_NMAT_VTI * N + mem) + pstd::grid_bytes(nx2, nz2, np, e) +
           e * nrec * max_steps + 12 * nrec;
}

int sm_pve_create(sm_ctx *ctx, const sm_pve_params *pp, const double *K_m, const double *K_s, const double *K_f,
                  const double *phi, const double *mu, const double *rho_s, const double *rho_f, const double *T,
                  const double *eta, const double *kappa, const double *Q, const double *fq, sm_pve **out)
{
    const PveModel in{K_m, K_s, K_f, phi, mu, rho_s, rho_f, T, T, eta, kappa, kappa, Q, fq, nullptr, nullptr};
    return create(ctx, pp, false, in, out);
}

int sm_pve_create_vti(sm_ctx *ctx, const sm_pve_params *pp, const double *K_m, const double *K_s, const double *K_f,
                      const double *phi, const double *mu, const double *rho_s, const double *rho_f, const double *T1,
                      const double *T3, const double *eta, const double *kappa1, const double *kappa3, const double *Q,
                      const double *fq, const double *epsilon, const double *delta, sm_pve **out)
{
    const PveModel in{K_m, K_s, K_f, phi, mu, rho_s, rho_f, T1, T3, eta, kappa1, kappa3, Q, fq, epsilon, delta};
    return create(ctx, pp, true, in, out);
}

int sm_pve_destroy(sm_pve *s)
{
    if (!s) return SM_OK;
    (void)set_device(s->ctx);
    (void)hipStreamSynchronize(s->ctx->stream);
    release(s);
    delete s;
    return SM_OK;
}

int sm_pve_set_sources(sm_pve *s, size_t nsrc, const int *type, const size_t *idx, const size_t *length,
                       const double *fct)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_pve_set_sources: NULL solver");
    SM_CHECK(nsrc == 0 || (type && idx && length && fct), SM_ERR_ARG, "sm_pve_set_sources: NULL array");
    SM_CHECK(nsrc < (1u << 20), SM_ERR_ARG, "sm_pve_set_sources: too many sources");
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

int sm_pve_set_receivers(sm_pve *s, size_t nrec, const size_t *idx, const int *comp, size_t max_steps)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_pve_set_receivers: NULL solver");
    SM_CHECK(nrec == 0 || (idx && comp), SM_ERR_ARG, "sm_pve_set_receivers: NULL array");
    SM_CHECK(nrec < (1u << 24), SM_ERR_ARG, "sm_pve_set_receivers: too many receivers");
    bool div = false, curl = false;
    for (size_t r = 0; r < nrec; ++r) {
        SM_CHECK(idx[r] < s->N, SM_ERR_ARG, "receiver %zu: index %zu outside the grid", r, idx[r]);
        SM_CHECK(comp_slot(comp[r]) >= 0 || comp[r] == SM_DIV || comp[r] == SM_CURL, SM_ERR_ARG,
                 "receiver %zu: component %d not recorded by the poroelastic solver", r, comp[r]);
        div = div || comp[r] == SM_DIV;
        curl = curl || comp[r] == SM_CURL;
    }
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    dfree(s->rcv); dfree(s->trace);
    s->nrec = nrec; s->max_steps = max_steps; s->rcv_step0 = s->step;
    s->need_div = div; s->need_curl = curl;
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

int sm_pve_set_energy_axes(sm_pve *s, size_t i1, size_t i2, size_t j1, size_t j2, const double *ax,
                           const double *bx, const double *cx, const double *az, const double *bz, const double *cz,
                           size_t max_steps)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_pve_set_energy: NULL solver");
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    for (sm_dbuf *b : {&s->ecoef, &s->epart, &s->etrace}) dfree(*b);
    s->e_max = 0;
    if (max_steps == 0) return SM_OK;
    SM_CHECK(ax && bx && cx && az && bz && cz, SM_ERR_ARG, "sm_pve_set_energy: NULL array");
    SM_CHECK(i1 <= i2 && i2 < (size_t)s->g.nx2 && j1 <= j2 && j2 < (size_t)s->g.nz2, SM_ERR_ARG,
             "sm_pve_set_energy: region [%zu, %zu] x [%zu, %zu] outside the %d x %d grid", i1, i2, j1, j2,
             s->g.nx2, s->g.nz2);
    const size_t ni = i2 - i1 + 1, nj = j2 - j1 + 1, nc = ni * nj;
    std::vector<double> h(6 * nc);
    const double *src[6] = {ax, bx, cx, az, bz, cz};
    for (int q = 0; q < 6; ++q) std::memcpy(h.data() + q * nc, src[q], nc * sizeof(double));
    int rc = dalloc(s->ecoef, h.size() * sizeof(double));
    if (rc == SM_OK) rc = dalloc(s->epart, E_BLOCKS * sizeof(double));
    if (rc == SM_OK) rc = dalloc(s->etrace, max_steps * sizeof(double));
    if (rc != SM_OK) return rc;
    SM_HIP(hipMemcpyAsync(s->ecoef.ptr, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s->ctx->stream));
    SM_HIP(hipMemsetAsync(s->etrace.ptr, 0, s->etrace.bytes, s->ctx->stream));
    SM_HIP(hipStreamSynchronize(s->ctx->stream));
    s->ei1 = i1; s->ej1 = j1; s->eni = ni; s->enj = nj;
    s->e_max = max_steps; s->e_step0 = s->step;
    return SM_OK;
}

int sm_pve_set_energy(sm_pve *s, size_t i1, size_t i2, size_t j1, size_t j2, const double *rho11,
                      const double *rho12, const double *rho22, size_t max_steps)
{
    return sm_pve_set_energy_axes(s, i1, i2, j1, j2, rho11, rho12, rho22, rho11, rho12, rho22, max_steps);
}

int sm_pve_get_energy(sm_pve *s, size_t it0, size_t count, double *out)
{
    SM_CHECK(s && (count == 0 || out), SM_ERR_ARG, "sm_pve_get_energy: NULL argument");
    SM_CHECK(it0 + count <= s->e_max, SM_ERR_ARG, "sm_pve_get_energy: rows [%zu, %zu) beyond %zu", it0,
             it0 + count, s->e_max);
    if (count == 0) return SM_OK;
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    SM_HIP(hipMemcpyAsync(out, (const double *)s->etrace.ptr + it0, count * sizeof(double), hipMemcpyDeviceToHost,
                          s->ctx->stream));
    SM_HIP(hipStreamSynchronize(s->ctx->stream));
    return SM_OK;
}

int sm_pve_run(sm_pve *s, size_t nsteps)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_pve_run: NULL solver");
    SM_CHECK(s->step + nsteps < (1ull << 61), SM_ERR_ARG, "sm_pve_run: step count too large");
    if (nsteps == 0) return SM_OK;
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    return s->elt == 8 ? run<double>(s, nsteps) : run<float>(s, nsteps);
}

// checkpoint blob: a header of 10 64-bit words (magic, element size, N,
// field values, memory values, trace values, energy rows, step, first
// trace step, first energy step), then the fields, the three memory-variable
// copies, the traces (element size) and the energy record (FP64)
namespace {
constexpr uint64_t CK_MAGIC = 0x31765045504D53ull;   // "SMPEPv1"
size_t ck_mem(const sm_pve *s) { return 3 * 6 * (s->mem_dx + s->mem_dz); }
}  // namespace

int sm_pve_state_bytes(const sm_pve *s, size_t *bytes)
{
    SM_CHECK(s && bytes, SM_ERR_ARG, "sm_pve_state_bytes: NULL argument");
    *bytes = 10 * sizeof(uint64_t) + s->elt * (W_NF * s->N + ck_mem(s) + s->nrec * s->max_steps) +
             sizeof(double) * s->e_max;
    return SM_OK;
}

int sm_pve_get_state(sm_pve *s, void *buf, size_t bytes)
{
    size_t need = 0;
    SM_CHECK(s && buf, SM_ERR_ARG, "sm_pve_get_state: NULL argument");
    sm_pve_state_bytes(s, &need);
    SM_CHECK(bytes >= need, SM_ERR_ARG, "sm_pve_get_state: %zu bytes, %zu needed", bytes, need);
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    uint64_t *h = (uint64_t *)buf;
    const uint64_t hdr[10] = {CK_MAGIC, s->elt, s->N, W_NF * s->N, ck_mem(s), s->nrec * s->max_steps, s->e_max,
                              s->step, s->rcv_step0, s->e_step0};
    std::memcpy(h, hdr, sizeof(hdr));
    char *p = (char *)buf + sizeof(hdr);
    hipStream_t st = s->ctx->stream;
    const size_t fw = W_NF * s->N * s->elt, mw = ck_mem(s) * s->elt, tw = s->nrec * s->max_steps * s->elt;
    SM_HIP(hipMemcpyAsync(p, (const char *)s->state.ptr + (size_t)s->iw * W_NF * s->N * s->elt, fw,
                          hipMemcpyDeviceToHost, st));
    SM_HIP(hipMemcpyAsync(p + fw, s->mem.ptr, mw, hipMemcpyDeviceToHost, st));
    if (tw) SM_HIP(hipMemcpyAsync(p + fw + mw, s->trace.ptr, tw, hipMemcpyDeviceToHost, st));
    if (s->e_max)
        SM_HIP(hipMemcpyAsync(p + fw + mw + tw, s->etrace.ptr, s->e_max * sizeof(double), hipMemcpyDeviceToHost, st));
    SM_HIP(hipStreamSynchronize(st));
    return SM_OK;
}

int sm_pve_set_state(sm_pve *s, const void *buf, size_t bytes)
{
    size_t need = 0;
    SM_CHECK(s && buf, SM_ERR_ARG, "sm_pve_set_state: NULL argument");
    sm_pve_state_bytes(s, &need);
    SM_CHECK(bytes >= need, SM_ERR_ARG, "sm_pve_set_state: %zu bytes, %zu needed", bytes, need);
    uint64_t h[10];
    std::memcpy(h, buf, sizeof(h));
    SM_CHECK(h[0] == CK_MAGIC && h[1] == s->elt && h[2] == s->N && h[3] == W_NF * s->N && h[4] == ck_mem(s) &&
                 h[5] == s->nrec * s->max_steps && h[6] == s->e_max,
             SM_ERR_ARG, "sm_pve_set_state: the state was saved from a solver of another shape, storage or record");
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    const char *p = (const char *)buf + sizeof(h);
    hipStream_t st = s->ctx->stream;
    const size_t fw = W_NF * s->N * s->elt, mw = ck_mem(s) * s->elt, tw = s->nrec * s->max_steps * s->elt;
    SM_HIP(hipMemcpyAsync((char *)s->state.ptr + (size_t)s->iw * W_NF * s->N * s->elt, p, fw, hipMemcpyHostToDevice,
                          st));
    SM_HIP(hipMemcpyAsync(s->mem.ptr, p + fw, mw, hipMemcpyHostToDevice, st));
    if (tw) SM_HIP(hipMemcpyAsync(s->trace.ptr, p + fw + mw, tw, hipMemcpyHostToDevice, st));
    if (s->e_max)
        SM_HIP(hipMemcpyAsync(s->etrace.ptr, p + fw + mw + tw, s->e_max * sizeof(double), hipMemcpyHostToDevice, st));
    SM_HIP(hipStreamSynchronize(st));
    s->step = h[7];
    s->rcv_step0 = h[8];
    s->e_step0 = h[9];
    return SM_OK;
}

int sm_pve_set_timing(sm_pve *s, int on)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_pve_set_timing: NULL solver");
    s->timing = on != 0;
    s->timed_ms = 0.0;
    s->timed_steps = 0;
    return SM_OK;
}

int sm_pve_get_timing(const sm_pve *s, double *ms_total, size_t *steps)
{
    SM_CHECK(s && ms_total && steps, SM_ERR_ARG, "sm_pve_get_timing: NULL argument");
    *ms_total = s->timed_ms;
    *steps = s->timed_steps;
    return SM_OK;
}

int sm_pve_current_step(const sm_pve *s, size_t *it)
{
    SM_CHECK(s && it, SM_ERR_ARG, "sm_pve_current_step: NULL argument");
    *it = s->step;
    return SM_OK;
}

int sm_pve_get_traces(sm_pve *s, size_t it0, size_t count, double *out)
{
    SM_CHECK(s && (count == 0 || out), SM_ERR_ARG, "sm_pve_get_traces: NULL argument");
    SM_CHECK(it0 + count <= s->max_steps, SM_ERR_ARG, "sm_pve_get_traces: rows [%zu, %zu) beyond %zu", it0,
             it0 + count, s->max_steps);
    if (count == 0 || s->nrec == 0) return SM_OK;
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    const char *base = (const char *)s->trace.ptr + it0 * s->nrec * s->elt;
    return s->elt == 8 ? read_back<double>(s, base, count * s->nrec, out)
                       : read_back<float>(s, base, count * s->nrec, out);
}

int sm_pve_get_field(sm_pve *s, int comp, double *out)
{
    SM_CHECK(s && out, SM_ERR_ARG, "sm_pve_get_field: NULL argument");
    const int q = comp_slot(comp);
    SM_CHECK(q >= 0 || comp == SM_DIV || comp == SM_CURL, SM_ERR_ARG, "sm_pve_get_field: component %d is not a field",
             comp);
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    const size_t N = s->N, e = s->elt;
    if (q >= 0) {
        const char *base = (const char *)s->state.ptr + ((size_t)s->iw * W_NF + q) * N * e;
        return e == 8 ? read_back<double>(s, base, N, out) : read_back<float>(s, base, N, out);
    }
    // divergence or curl of the last stage's input (zero before any step)
    const bool div = comp == SM_DIV;
    int rc = e == 8 ? div_curl<double>(s, STATE<double>(s, s->iwt), div, !div)
                    : div_curl<float>(s, STATE<float>(s, s->iwt), div, !div);
    if (rc != SM_OK) return rc;
    std::vector<double> a(N), b(N);
    const char *d = (const char *)s->der.ptr;
    const int sx = div ? D_DIVX : D_CURLX, sz = div ? D_DIVZ : D_CURLZ;
    rc = e == 8 ? read_back<double>(s, d + sx * N * e, N, a.data()) : read_back<float>(s, d + sx * N * e, N, a.data());
    if (rc == SM_OK)
        rc = e == 8 ? read_back<double>(s, d + sz * N * e, N, b.data()) : read_back<float>(s, d + sz * N * e, N, b.data());
    if (rc != SM_OK) return rc;
    for (size_t n = 0; n < N; ++n) out[n] = div ? a[n] + b[n] : a[n] - b[n];
    return SM_OK;
}

int sm_pve_reset(sm_pve *s)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_pve_reset: NULL solver");
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    s->step = 0;
    s->rcv_step0 = 0;
    s->e_step0 = 0;
    return zero_state(s);
}

int sm_pve_device_bytes(const sm_pve *s, size_t *bytes)
{
    SM_CHECK(s && bytes, SM_ERR_ARG, "sm_pve_device_bytes: NULL argument");
    size_t b = pstd::grid_device_bytes(s->g);
    for (const sm_dbuf *d : {&s->state, &s->der, &s->mat, &s->mem, &s->src, &s->rcv, &s->trace, &s->ecoef,
                             &s->epart, &s->etrace})
        b += d->bytes;
    *bytes = b;
    return SM_OK;
}

}  // extern "C"
