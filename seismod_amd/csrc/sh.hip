// sh.hip -- viscoelastic SH waves in a 2D VTI medium (ve_vti_sh_pml):
// pseudo-spectral derivatives on a staggered grid, CPML, fourth-order
// Runge-Kutta in time, one standard-linear-solid relaxation per stress.
//
// The reference (ve_vti_sh_pml.c:613-700, propagate.c:2280-2462) evaluates
// four derivative stages per step; each stage is three FFT derivatives
// (FFTW r2c, multiply by i k e^{+-i k d/2}, c2r), the CPML corrections of
// their boundary layers, and the pointwise right-hand sides.  Here every
// FFT is a batched hipFFT plan over the padded nx2 x nz2 grid (z fastest):
// along x with stride nz2, along z with unit stride, FP64 or FP32.  The
// k multipliers fold the 1/n of the unnormalised inverse, as the reference
// does.  The CPML layers are one launch per memory-variable set, the
// index of every slot being the reference's region-loop counter; the
// right-hand sides, the Runge-Kutta combinations and the source injection
// are elementwise kernels.  Results match the FP64 restatement to FFT
// rounding (the FFT libraries differ), not bit for bit.
//
// Deviations (undefined reads in the reference, given their natural
// values): c44_0 of the last x row uses that row's filled eta2; the tables
// of the right-side dx_txy loop at nl = np (one past their end) take their
// own formula there (k = 1, c = 0); dx_vy, indexed like a z layer, gets an
// array large enough for nx2 columns.
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "sm_internal.hpp"

namespace sm {
namespace {

enum { W_TXY, W_TYZ, W_VY, W_E1, W_E2, W_NF };
enum { M_C44, M_C66, M_C440, M_C660, M_TS1, M_TS2, M_ETA1, M_ETA2, M_IRHO, M_NMAT };
// CPML table rows (np + 1 entries each): per stage b, c, bh, ch, bH, cH for
// x then z; shared k, kh, kH for x then z
enum { T_B, T_C, T_BH, T_CH, T_BHE, T_CHE, T_NST };
enum { K_K, K_KH, K_KHE, K_NK };

template <typename T>
struct Cx {
    T re, im;
};

template <typename T>
__global__ void sh_kmul(Cx<T> *o, const Cx<T> *k, int nbins, size_t n)
{
    const size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const Cx<T> a = o[l], b = k[l % nbins];
    o[l] = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// one CPML memory-variable set: slot q of `count` maps to a cell and a
// table row exactly as the reference's loop counter ipml does.
//   mode 0 (x layers over nz2 columns): q < np*nz2 -> row nl = q / nz2 at
//     cell (nl, j); then rows nl <= lastn at cell (nx2-1-nl, j)
//   mode 1 (z layers over nx2 columns): q < np*nx2 -> nl = q / nx2 at cell
//     (i, nl); then nl <= lastn at cell (i, nz2-1-nl)
template <typename T>
__global__ void sh_cpml(T *t, T *m1, const T *m0, const T *m2, int mode, int np, int lastn, int nx2,
                        int nz2, const T *b1, const T *c1, const T *k1, const T *b2, const T *c2,
                        const T *k2)
{
    const int cols = mode == 0 ? nz2 : nx2;
    const size_t count = (size_t)(np + lastn + 1) * cols;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const bool first = q < (size_t)np * cols;
    const size_t r = first ? q : q - (size_t)np * cols;
    const int nl = (int)(r / cols), col = (int)(r % cols);
    size_t ind;
    if (mode == 0) ind = (size_t)(first ? nl : nx2 - 1 - nl) * nz2 + col;
    else ind = (size_t)col * nz2 + (first ? nl : nz2 - 1 - nl);
    const T b = first ? b1[nl] : b2[nl], c = first ? c1[nl] : c2[nl], k = first ? k1[nl] : k2[nl];
    const T d = t[ind];
    m1[q] = b * m0[q] + c * d;
    t[ind] = d / k + m2[q];
}

// t1 += t2; v2_y = (1/rho) t1 on the interior nodes (propagate.c:2364-2374)
template <typename T>
__global__ void sh_rhs_v(const T *t1, const T *t2, const T *irho, T *dv, int nx2, int nz2)
{
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)nx2 * nz2) return;
    const int i = (int)(n / nz2), j = (int)(n % nz2);
    if (i < 1 || i >= nx2 - 1 || j < 1 || j >= nz2 - 1) return;
    dv[n] = irho[n] * (t1[n] + t2[n]);
}

// e' = (1/tau_s)((1 - 1/eta) D v - e);  tau' = c D v + c_0 e
template <typename T>
__global__ void sh_rhs_s(const T *t1, const T *ts, const T *eta, const T *c, const T *c0, const T *e,
                         T *de, T *dtau, size_t n)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const T d = t1[q], ee = e[q];
    de[q] = ts[q] * (eta[q] * d - ee);
    dtau[q] = c[q] * d + c0[q] * ee;
}

// Runge-Kutta combinations over the 5 N state values:
//   wt = ws + a delta (when set_t), w = (acc ? w : ws) + b delta
template <typename T>
__global__ void sh_rk(T *w, T *wt, const T *ws, const T *delta, T a, T b, int set_t, int acc, size_t n)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const T d = delta[q];
    if (set_t) wt[q] = ws[q] + a * d;
    w[q] = (acc ? w[q] : ws[q]) + b * d;
}

template <typename T>
__global__ void sh_copy(T *dst, const T *src, size_t n)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) dst[q] = src[q];
}

// add_src_ve_sh_rk4 (src.c:145-168), in source order on one thread
template <typename T>
__global__ void sh_src(T *w, T *wt, size_t N, int nsrc, const int64_t *tab, const T *fct, T dtw,
                       T dtt, int64_t it)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int64_t k = 2 * it + 1;
    for (int s = 0; s < nsrc; ++s) {
        const int64_t typ = tab[s], cell = tab[nsrc + s], off = tab[2 * nsrc + s], len = tab[3 * nsrc + s];
        if (k >= len) continue;
        const size_t q = typ == SM_SXY ? 0 : typ == SM_SYZ ? N : ~(size_t)0;
        if (q == ~(size_t)0) continue;
        const T v = fct[off + k];
        if (dtt != T(0)) wt[q + cell] += dtt * v;
        w[q + cell] += dtw * v;
    }
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
    int nx2 = 0, nz2 = 0;
    size_t N = 0, elt = 8;
    int nxh = 0, nzh = 0;           // rfft bins along x and z
    // W, Ws, Wt, delta (5 N each), t1, t2 (N), material (9 N)
    sm_dbuf state, tmp, mat;
    sm_dbuf spec;                   // o_x ((nx2/2+1) nz2) then o_z (nx2 (nz2/2+1)), complex
    sm_dbuf kmul;                   // kx_f, kx_b, kz_f, kz_b
    sm_dbuf tabs;                   // per stage 2 x 6 rows, then 2 x 3 shared k rows (np + 1 each)
    sm_dbuf mem;                    // 3 copies x (dx_txy, dx_vy, dz_tyz, dz_vy)
    size_t mem_dx = 0, mem_dxv = 0, mem_dz = 0;   // entries of dx_txy, dx_vy, dz_* per copy
    sm_dbuf src, rcv, trace;
    size_t nsrc = 0, nrec = 0, max_steps = 0, rcv_step0 = 0, step = 0;
    hipfftHandle px_f = 0, px_i = 0, pz_f = 0, pz_i = 0;
    bool timing = false;
    double timed_ms = 0.0;
    size_t timed_steps = 0;
};

namespace {

void release(sm_sh *s)
{
    for (hipfftHandle *h : {&s->px_f, &s->px_i, &s->pz_f, &s->pz_i})
        if (*h) { (void)hipfftDestroy(*h); *h = 0; }
    for (sm_dbuf *b : {&s->state, &s->tmp, &s->mat, &s->spec, &s->kmul, &s->tabs, &s->mem, &s->src,
                       &s->rcv, &s->trace})
        dfree(*b);
}

size_t mem_entries(size_t np, size_t nx2, size_t nz2, size_t *dx, size_t *dxv, size_t *dz)
{
    *dx = (2 * np + 1) * nz2;
    *dxv = (2 * np + 1) * std::max(nx2, nz2);
    *dz = (2 * np + 1) * nx2;
    return 3 * (*dx + *dxv + 2 * *dz);
}

template <typename T>
size_t tab_entries(size_t np)
{
    return (3 * 2 * T_NST + 2 * K_NK) * (np + 1);
}

// device views
template <typename T> T *W(sm_sh *s, int which)   // 0 W, 1 Ws, 2 Wt, 3 delta
{
    return (T *)s->state.ptr + (size_t)which * W_NF * s->N;
}
template <typename T> T *TMP(sm_sh *s, int q) { return (T *)s->tmp.ptr + (size_t)q * s->N; }
template <typename T> const T *MAT(sm_sh *s, int q) { return (const T *)s->mat.ptr + (size_t)q * s->N; }
template <typename T> T *MEM(sm_sh *s, int copy, int set)   // set 0 dx_txy, 1 dx_vy, 2 dz_tyz, 3 dz_vy
{
    const size_t per = s->mem_dx + s->mem_dxv + 2 * s->mem_dz;
    const size_t off[4] = {0, s->mem_dx, s->mem_dx + s->mem_dxv, s->mem_dx + s->mem_dxv + s->mem_dz};
    return (T *)s->mem.ptr + copy * per + off[set];
}
// stage 0 (fp1), 1 (fp23), 2 (fp4); axis 0 x, 1 z
template <typename T> const T *TAB(sm_sh *s, int stage, int axis, int row)
{
    return (const T *)s->tabs.ptr + ((size_t)(stage * 2 + axis) * T_NST + row) * (s->p.np + 1);
}
template <typename T> const T *KTAB(sm_sh *s, int axis, int row)
{
    return (const T *)s->tabs.ptr + ((size_t)3 * 2 * T_NST + axis * K_NK + row) * (s->p.np + 1);
}

inline unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

hipfftResult exec_f(hipfftHandle p, const double *in, void *out)
{
    return hipfftExecD2Z(p, const_cast<double *>(in), (hipfftDoubleComplex *)out);
}
hipfftResult exec_f(hipfftHandle p, const float *in, void *out)
{
    return hipfftExecR2C(p, const_cast<float *>(in), (hipfftComplex *)out);
}
hipfftResult exec_i(hipfftHandle p, void *in, double *out)
{
    return hipfftExecZ2D(p, (hipfftDoubleComplex *)in, out);
}
hipfftResult exec_i(hipfftHandle p, void *in, float *out)
{
    return hipfftExecC2R(p, (hipfftComplex *)in, out);
}

#define SH_FFT(call)                                                                     \
    do {                                                                                 \
        hipfftResult r_ = (call);                                                        \
        if (r_ != HIPFFT_SUCCESS) return fail(SM_ERR_HIP, "%s: hipFFT error %d", #call, (int)r_); \
    } while (0)

// derivative of field f along axis (0 x, 1 z), forward (D+) or backward
// (D-), into out
template <typename T>
int deriv(sm_sh *s, const T *f, int axis, bool fwd, T *out)
{
    hipStream_t st = s->ctx->stream;
    Cx<T> *spec = (Cx<T> *)s->spec.ptr;
    Cx<T> *ox = spec, *oz = spec + (size_t)s->nxh * s->nz2;
    const Cx<T> *km = (const Cx<T> *)s->kmul.ptr;
    if (axis == 0) {
        SH_FFT(exec_f(s->px_f, f, ox));
        const size_t n = (size_t)s->nxh * s->nz2;
        sh_kmul<T><<<blocks(n), 256, 0, st>>>(ox, km + (fwd ? 0 : s->nxh), s->nxh, n);
        SH_FFT(exec_i(s->px_i, ox, out));
    } else {
        SH_FFT(exec_f(s->pz_f, f, oz));
        const size_t n = (size_t)s->nzh * s->nx2;
        sh_kmul<T><<<blocks(n), 256, 0, st>>>(oz, km + 2 * s->nxh + (fwd ? 0 : s->nzh), s->nzh, n);
        SH_FFT(exec_i(s->pz_i, oz, out));
    }
    return SM_OK;
}

template <typename T>
void cpml(sm_sh *s, T *t, int set, int mode, int lastn, const T *b1, const T *c1, const T *k1,
          const T *b2, const T *c2, const T *k2)
{
    const int np = (int)s->p.np;
    const int cols = mode == 0 ? s->nz2 : s->nx2;
    const size_t count = (size_t)(np + lastn + 1) * cols;
    if (count == 0) return;   // no layers at all (dz_tyz without abs)
    sh_cpml<T><<<blocks(count), 256, 0, s->ctx->stream>>>(t, MEM<T>(s, 1, set), MEM<T>(s, 0, set),
                                                          MEM<T>(s, 2, set), mode, np, lastn, s->nx2,
                                                          s->nz2, b1, c1, k1, b2, c2, k2);
}

// propagateVE_SH_CPML on Wt -> delta with the tables of `stage`
template <typename T>
int propagate(sm_sh *s, int stage)
{
    hipStream_t st = s->ctx->stream;
    const size_t N = s->N;
    const int np = (int)s->p.np;
    T *wt = W<T>(s, 2), *dl = W<T>(s, 3);
    T *t1 = TMP<T>(s, 0), *t2 = TMP<T>(s, 1);
    int rc;
    if ((rc = deriv<T>(s, wt + W_TXY * N, 0, false, t1)) != SM_OK) return rc;   // D_x^- tau_xy
    if ((rc = deriv<T>(s, wt + W_TYZ * N, 1, false, t2)) != SM_OK) return rc;   // D_z^- tau_yz
    // dx_txy: left rows nl < np, right rows nx2-1-nl for nl <= np
    cpml<T>(s, t1, 0, 0, np, TAB<T>(s, stage, 0, T_B), TAB<T>(s, stage, 0, T_C), KTAB<T>(s, 0, K_K),
            TAB<T>(s, stage, 0, T_B), TAB<T>(s, stage, 0, T_C), KTAB<T>(s, 0, K_K));
    // dz_tyz: top and bottom nl < np
    cpml<T>(s, t2, 2, 1, np - 1, TAB<T>(s, stage, 1, T_B), TAB<T>(s, stage, 1, T_C), KTAB<T>(s, 1, K_K),
            TAB<T>(s, stage, 1, T_B), TAB<T>(s, stage, 1, T_C), KTAB<T>(s, 1, K_K));
    sh_rhs_v<T><<<blocks(N), 256, 0, st>>>(t1, t2, MAT<T>(s, M_IRHO), dl + W_VY * N, s->nx2, s->nz2);
    // D_z^+ v_y; top half nodes, bottom end half nodes nl <= np
    if ((rc = deriv<T>(s, wt + W_VY * N, 1, true, t1)) != SM_OK) return rc;
    cpml<T>(s, t1, 3, 1, np, TAB<T>(s, stage, 1, T_BH), TAB<T>(s, stage, 1, T_CH), KTAB<T>(s, 1, K_KH),
            TAB<T>(s, stage, 1, T_BHE), TAB<T>(s, stage, 1, T_CHE), KTAB<T>(s, 1, K_KHE));
    sh_rhs_s<T><<<blocks(N), 256, 0, st>>>(t1, MAT<T>(s, M_TS1), MAT<T>(s, M_ETA1), MAT<T>(s, M_C44),
                                           MAT<T>(s, M_C440), wt + W_E1 * N, dl + W_E1 * N,
                                           dl + W_TYZ * N, N);
    // D_x^+ v_y; its layers indexed like z layers, x tables on top and the
    // z end tables below, as the reference has them
    if ((rc = deriv<T>(s, wt + W_VY * N, 0, true, t1)) != SM_OK) return rc;
    cpml<T>(s, t1, 1, 1, np, TAB<T>(s, stage, 0, T_BH), TAB<T>(s, stage, 0, T_CH), KTAB<T>(s, 0, K_KH),
            TAB<T>(s, stage, 1, T_BHE), TAB<T>(s, stage, 1, T_CHE), KTAB<T>(s, 1, K_KHE));
    sh_rhs_s<T><<<blocks(N), 256, 0, st>>>(t1, MAT<T>(s, M_TS2), MAT<T>(s, M_ETA2), MAT<T>(s, M_C66),
                                           MAT<T>(s, M_C660), wt + W_E2 * N, dl + W_E2 * N,
                                           dl + W_TXY * N, N);
    // m2 = m1 over the reference's ranges (propagate.c:2451-2461)
    const size_t n1 = (size_t)(2 * np + 1) * s->nz2, n2 = (size_t)(2 * np + 1) * s->nx2;
    sh_copy<T><<<blocks(n1), 256, 0, st>>>(MEM<T>(s, 2, 0), MEM<T>(s, 1, 0), n1);
    sh_copy<T><<<blocks(n1), 256, 0, st>>>(MEM<T>(s, 2, 1), MEM<T>(s, 1, 1), n1);
    sh_copy<T><<<blocks(n2), 256, 0, st>>>(MEM<T>(s, 2, 2), MEM<T>(s, 1, 2), n2);
    sh_copy<T><<<blocks(n2), 256, 0, st>>>(MEM<T>(s, 2, 3), MEM<T>(s, 1, 3), n2);
    return SM_OK;
}

template <typename T>
int run(sm_sh *s, size_t nsteps)
{
    hipStream_t st = s->ctx->stream;
    const size_t N = s->N, n5 = W_NF * N;
    const int np = (int)s->p.np;
    const T dt = (T)s->p.dt;
    T *w = W<T>(s, 0), *ws = W<T>(s, 1), *wt = W<T>(s, 2), *dl = W<T>(s, 3);
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
    for (size_t k = 0; k < nsteps; ++k) {
        const int64_t it = (int64_t)(s->step + k);
        if (s->timing) SM_HIP(hipEventRecord(ev[2 * k], st));
        sh_copy<T><<<blocks(n5), 256, 0, st>>>(ws, w, n5);
        const size_t n0x = (size_t)2 * np * s->nz2, n0z = (size_t)2 * np * s->nx2;
        if (n0x) {
            sh_copy<T><<<blocks(n0x), 256, 0, st>>>(MEM<T>(s, 0, 0), MEM<T>(s, 1, 0), n0x);
            sh_copy<T><<<blocks(n0x), 256, 0, st>>>(MEM<T>(s, 0, 1), MEM<T>(s, 1, 1), n0x);
            sh_copy<T><<<blocks(n0z), 256, 0, st>>>(MEM<T>(s, 0, 2), MEM<T>(s, 1, 2), n0z);
            sh_copy<T><<<blocks(n0z), 256, 0, st>>>(MEM<T>(s, 0, 3), MEM<T>(s, 1, 3), n0z);
        }
        sh_copy<T><<<blocks(n5), 256, 0, st>>>(wt, ws, n5);
        // eq. (80)-(83) and (79) of the reference's scheme
        const T h6 = dt / T(6), h3 = dt / T(3), h2 = dt / T(2);
        struct Stage { int tabs; T a; T b; int acc; T sw; T st; };
        const Stage stages[4] = {{0, h2, h6, 0, h6, h2}, {1, h2, h3, 1, h3, h2},
                                 {1, dt, h3, 1, h3, dt}, {2, T(0), h6, 1, h6, T(0)}};
        for (int q = 0; q < 4; ++q) {
            int rc = propagate<T>(s, stages[q].tabs);
            if (rc != SM_OK) return rc;
            sh_rk<T><<<blocks(n5), 256, 0, st>>>(w, wt, ws, dl, stages[q].a, stages[q].b, q < 3 ? 1 : 0,
                                                 stages[q].acc, n5);
            if (s->nsrc)
                sh_src<T><<<1, 64, 0, st>>>(w, wt, N, (int)s->nsrc, stab, sfct, stages[q].sw, stages[q].st,
                                            it);
        }
        if (s->timing) SM_HIP(hipEventRecord(ev[2 * k + 1], st));
        if (s->nrec && (size_t)it >= s->rcv_step0 && (size_t)it - s->rcv_step0 < s->max_steps)
            sh_gather<T><<<blocks(s->nrec), 256, 0, st>>>(w, N, (int)s->nrec, ridx, rcomp, (T *)s->trace.ptr,
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

// CPML tables of pml.c:1959-2106 for one axis, rows of np + 1 entries
// (entry np: the tables' own formulas there); k rows shared by the stages
void cpml_axis(const sm_sh_params &p, double d, double alpha_max, std::vector<double> st[3][T_NST],
               std::vector<double> kk[K_NK])
{
    const size_t n = p.np;
    for (int s = 0; s < 3; ++s)
        for (int r = 0; r < T_NST; ++r) st[s][r].assign(n + 1, 0.0);
    for (int r = 0; r < K_NK; ++r) kk[r].assign(n + 1, 1.0);
    if (!(p.abs && n > 0)) return;
    const double L = n * d;
    const double d0 = -(p.pml_order + 1.) * p.vmax * std::log(p.Rc) / (2. * L);
    const double da = alpha_max / n;
    const double theta = 0.5;
    auto bc = [&](double k, double dv, double a, double h, double &b, double &c) {
        b = (k - (1. - theta) * h * (dv + k * a)) / (k + theta * h * (dv + k * a));
        c = -dv * h / (k * (k + theta * h * (dv + k * a)));
    };
    for (size_t m = 0; m <= n; ++m) {
        const double x = (double)(n - m) * d / L;
        const double k = 1. + (p.kappa_max - 1.) * std::pow(x, p.pml_order);
        const double dv = d0 * std::pow(x, p.pml_order);
        const double al = m * da;
        const bool inner = m < n;
        const double xh = ((double)n - m - 0.5) * d / L;
        const double kh = inner ? 1. + (p.kappa_max - 1.) * std::pow(xh, p.pml_order) : 1.;
        const double dh = inner ? d0 * std::pow(xh, p.pml_order) : 0.;
        const double alh = (m + 0.5) * da;
        const double xH = ((double)n - m + 0.5) * d / L;
        const double kH = 1. + (p.kappa_max - 1.) * std::pow(xH, p.pml_order);
        const double dH = d0 * std::pow(xH, p.pml_order);
        const double alH = (m - 0.5) * da;
        kk[K_K][m] = k; kk[K_KH][m] = kh; kk[K_KHE][m] = kH;
        st[0][T_B][m] = st[0][T_BH][m] = st[0][T_BHE][m] = 1.0;
        for (int s = 1; s < 3; ++s) {
            const double h = s == 1 ? p.dt * 0.5 : p.dt;
            bc(k, dv, al, h, st[s][T_B][m], st[s][T_C][m]);
            bc(kh, dh, alh, h, st[s][T_BH][m], st[s][T_CH][m]);
            bc(kH, dH, alH, h, st[s][T_BHE][m], st[s][T_CHE][m]);
        }
    }
}

template <typename T>
int upload(void *dst, const std::vector<double> &v, hipStream_t st)
{
    std::vector<T> t(v.begin(), v.end());
    SM_HIP(hipMemcpyAsync(dst, t.data(), t.size() * sizeof(T), hipMemcpyHostToDevice, st));
    SM_HIP(hipStreamSynchronize(st));
    return SM_OK;
}

int make_plan(hipfftHandle *h, int n, int howmany, int istride, int idist, int ostride, int odist,
              int in_total, int out_total, hipfftType type, hipStream_t st)
{
    int nn[1] = {n}, inembed[1] = {in_total}, onembed[1] = {out_total};
    hipfftResult r = hipfftPlanMany(h, 1, nn, inembed, istride, idist, onembed, ostride, odist, type, howmany);
    if (r != HIPFFT_SUCCESS) return fail(SM_ERR_HIP, "hipfftPlanMany (n %d, batch %d): error %d", n, howmany, (int)r);
    r = hipfftSetStream(*h, st);
    if (r != HIPFFT_SUCCESS) return fail(SM_ERR_HIP, "hipfftSetStream: error %d", (int)r);
    return SM_OK;
}

int zero_state(sm_sh *s)
{
    hipStream_t st = s->ctx->stream;
    SM_HIP(hipMemsetAsync(s->state.ptr, 0, s->state.bytes, st));
    SM_HIP(hipMemsetAsync(s->mem.ptr, 0, s->mem.bytes, st));
    if (s->trace.ptr) SM_HIP(hipMemsetAsync(s->trace.ptr, 0, s->trace.bytes, st));
    SM_HIP(hipStreamSynchronize(st));
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
    const size_t spec = ((nx2 / 2 + 1) * nz2 + nx2 * (nz2 / 2 + 1)) * 2 * e;
    return e * (4 * W_NF * N + 2 * N + M_NMAT * N + mem) + spec + 2 * e * (nx2 + nz2 + 2) * 2 +
           e * (3 * 2 * T_NST + 2 * K_NK) * (np + 1) + e * nrec * max_steps + 12 * nrec;
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
    SM_CHECK(p.flags == 0, SM_ERR_ARG, "sm_sh_create: unknown flags %#x", p.flags);
    if (!p.abs) p.np = 0;   // io_utils.c:293: no absorbing layers without abs
    const size_t nx2 = p.nx + 2 * p.np, nz2 = p.nz + 2 * p.np, N = nx2 * nz2;
    SM_CHECK(nx2 < (1u << 20) && nz2 < (1u << 20) && N < ((size_t)1 << 31), SM_ERR_ARG,
             "sm_sh_create: grid too large");
    for (size_t n = 0; n < N; ++n)
        SM_CHECK(rho[n] > 0 && std::isfinite(mu[n]) && std::isfinite(gamma[n]) && Q[n] > 0 && f[n] > 0,
                 SM_ERR_ARG, "sm_sh_create: cell %zu needs rho, Q, f > 0 and finite mu, gamma", n);
    SM_CHECK(set_device(ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");

    sm_sh *s = new sm_sh;
    s->ctx = ctx;
    s->p = p;
    s->nx2 = (int)nx2; s->nz2 = (int)nz2; s->N = N;
    s->elt = p.storage == SM_STORE_F64 ? 8 : 4;
    s->nxh = (int)(nx2 / 2 + 1); s->nzh = (int)(nz2 / 2 + 1);
    const size_t e = s->elt;
    const size_t mem = mem_entries(p.np, nx2, nz2, &s->mem_dx, &s->mem_dxv, &s->mem_dz);
    int rc = dalloc(s->state, 4 * W_NF * N * e);
    if (rc == SM_OK) rc = dalloc(s->tmp, 2 * N * e);
    if (rc == SM_OK) rc = dalloc(s->mat, M_NMAT * N * e);
    if (rc == SM_OK) rc = dalloc(s->spec, ((size_t)s->nxh * nz2 + nx2 * (size_t)s->nzh) * 2 * e);
    if (rc == SM_OK) rc = dalloc(s->kmul, 2 * ((size_t)s->nxh + s->nzh) * 2 * e);
    if (rc == SM_OK) rc = dalloc(s->tabs, (3 * 2 * T_NST + 2 * K_NK) * (p.np + 1) * e);
    if (rc == SM_OK) rc = dalloc(s->mem, mem * e);
    if (rc != SM_OK) { release(s); delete s; return rc; }
    hipStream_t st = ctx->stream;

    // material, ve_vti_sh_pml.c:475-585 (the last row's c44_0 uses the
    // filled eta2, see the file comment)
    {
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
        std::vector<double> all;
        all.reserve(M_NMAT * N);
        for (auto &a : m) all.insert(all.end(), a.begin(), a.end());
        for (double v : all)
            if (!std::isfinite(v)) { release(s); delete s; return fail(SM_ERR_ARG, "sm_sh_create: material not finite"); }
        rc = e == 8 ? upload<double>(s->mat.ptr, all, st) : upload<float>(s->mat.ptr, all, st);
    }
    // k multipliers i k e^{+-i k d/2}, k = 2 pi n / (d n2^2) (propagate.c:566-577)
    if (rc == SM_OK) {
        std::vector<double> km;
        auto add = [&](size_t n2, double d, double sgn) {
            const double dn = 8.0 * std::atan(1.0) / (d * n2 * n2);
            for (size_t n = 0; n < n2 / 2 + 1; ++n) {
                const double k = n * dn, ph = sgn * 0.5 * k * d * n2;
                km.push_back(-k * std::sin(ph));   // re(i k e^{i ph})
                km.push_back(k * std::cos(ph));    // im
            }
        };
        add(nx2, p.dx, 1.0); add(nx2, p.dx, -1.0); add(nz2, p.dz, 1.0); add(nz2, p.dz, -1.0);
        rc = e == 8 ? upload<double>(s->kmul.ptr, km, st) : upload<float>(s->kmul.ptr, km, st);
    }
    // CPML tables
    if (rc == SM_OK) {
        const double pi = 4.0 * std::atan(1.0);
        const double alpha_max = p.alpha_max ? pi * p.f0 : 0.0;
        std::vector<double> sx[3][T_NST], sz[3][T_NST], kx[K_NK], kz[K_NK];
        cpml_axis(p, p.dx, alpha_max, sx, kx);
        cpml_axis(p, p.dz, alpha_max, sz, kz);
        std::vector<double> t;
        for (int s3 = 0; s3 < 3; ++s3) {
            for (int r = 0; r < T_NST; ++r) t.insert(t.end(), sx[s3][r].begin(), sx[s3][r].end());
            for (int r = 0; r < T_NST; ++r) t.insert(t.end(), sz[s3][r].begin(), sz[s3][r].end());
        }
        for (int r = 0; r < K_NK; ++r) t.insert(t.end(), kx[r].begin(), kx[r].end());
        for (int r = 0; r < K_NK; ++r) t.insert(t.end(), kz[r].begin(), kz[r].end());
        rc = e == 8 ? upload<double>(s->tabs.ptr, t, st) : upload<float>(s->tabs.ptr, t, st);
    }
    // FFT plans: x transforms of stride nz2 into [nz2][nx2/2+1], z transforms
    // of unit stride into [nx2][nz2/2+1] (propagate.c:593-669)
    if (rc == SM_OK) {
        const hipfftType f_t = e == 8 ? HIPFFT_D2Z : HIPFFT_R2C, i_t = e == 8 ? HIPFFT_Z2D : HIPFFT_C2R;
        const int NX = (int)nx2, NZ = (int)nz2, NN = (int)N;
        rc = make_plan(&s->px_f, NX, NZ, NZ, 1, 1, s->nxh, NN, s->nxh * NZ, f_t, st);
        if (rc == SM_OK) rc = make_plan(&s->px_i, NX, NZ, 1, s->nxh, NZ, 1, s->nxh * NZ, NN, i_t, st);
        if (rc == SM_OK) rc = make_plan(&s->pz_f, NZ, NX, 1, NZ, 1, s->nzh, NN, NX * s->nzh, f_t, st);
        if (rc == SM_OK) rc = make_plan(&s->pz_i, NZ, NX, 1, s->nzh, 1, NZ, NX * s->nzh, NN, i_t, st);
    }
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
    const char *base = (const char *)s->state.ptr + (size_t)q * s->N * s->elt;
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
    size_t b = 0;
    for (const sm_dbuf *d : {&s->state, &s->tmp, &s->mat, &s->spec, &s->kmul, &s->tabs, &s->mem, &s->src,
                             &s->rcv, &s->trace})
        b += d->bytes;
    *bytes = b;
    return SM_OK;
}

}  // extern "C"
