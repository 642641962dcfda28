// pstd.hpp -- the pseudo-spectral machinery shared by the viscoelastic
// PSTD solvers (ve_vti_sh_pml, ve_vti_pml): staggered FFT derivatives on
// the padded nx2 x nz2 grid (z fastest), the CPML stage tables and the
// layer corrections of the derivatives, applied inside each solver's fused
// stage kernel.
//
// Derivatives (propagate.c:2463-2510): f -> IFFT(H(k) FFT(f)) along one
// axis, H(k) = i k e^{+-i k d / 2} with the 1/n of the unnormalised inverse
// folded into k (propagate.c:363-374, 566-577).
//   x (stride nz2): when nz2 is even, the real field is read as an
//     nx2 x nz2/2 array of complex numbers, z_j = f(., 2j) + i f(., 2j+1);
//     one complex FFT along x transforms two real columns at once, the
//     multiplier is extended to the whole spectrum (H(n - k) = conj H(k),
//     the real part of H at the Nyquist bin, as a complex-to-real inverse
//     takes it), and the inverse complex FFT returns D f(., 2j) in the real
//     and D f(., 2j+1) in the imaginary part.  No real/complex copy passes,
//     half the transforms.  Odd nz2 falls back to real transforms.
//   z (unit stride): real-to-complex and complex-to-real transforms.
// Results equal the real-transform derivatives to rounding.
//
// When an axis length factors into 2, 3, 5, 7, 11 and 13, its derivatives
// skip hipFFT: one pstd_dfft launch takes every derivative of a stage along
// that axis.  A workgroup loads a tile of row pairs (z) or column pairs (x)
// into LDS as complex numbers, runs a mixed-radix Stockham FFT in LDS,
// multiplies by the whole-spectrum multiplier, runs the inverse as the
// conjugate of the forward transform of the conjugate, and writes the two
// derivatives: one HBM read and one write per value instead of six
// (forward transform, multiplier pass, inverse transform).
//
// CPML (pml.c:1959-2106): per stage (fp1, fp23, fp4) and axis the rows b,
// c at nodes, bh, ch at inner half nodes, bH, cH at outer half nodes; the
// k rows are shared by the stages.  Every row has np + 1 entries; entry np
// of the node and inner-half rows (read by loops that run to nl = np) takes
// the row's own formula there (k = 1, d = 0: b = 1, c = 0 for the node row).
//
// Layer sets: one memory variable per slot q, mapped to a cell as the
// reference's region loops map their counter ipml:
//   mode 0 (x layers, nz2 columns): q < np nz2 -> row nl = q / nz2 at cell
//     (nl, j); then rows nl <= lastn at cell (nx2 - 1 - nl, j);
//   mode 1 (z layers, nx2 columns): q < np nx2 -> nl = q / nx2 at (i, nl);
//     then rows nl <= lastn at (i, nz2 - 1 - nl).
// At a slot: m1 = b m0 + c d; d <- d / k + m2.  The copies the reference
// makes between propagations are folded in: m2 <- m1 for q < lim2 (after
// every propagation) and, in the last stage of a step, m0 <- m1 for
// q < lim0 (the copy at the start of the next step).
#pragma once

#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "sm_internal.hpp"

namespace sm {
namespace pstd {

template <typename T>
struct Cx {
    T re, im;
};

// CPML table rows per stage and axis, then shared k rows per axis
enum { T_B, T_C, T_BH, T_CH, T_BHE, T_CHE, T_NST };
enum { K_K, K_KH, K_KHE, K_NK };

struct Cpml {
    int abs = 0;
    double vmax = 0, pml_order = 2, Rc = 1e-5, kappa_max = 1, alpha_max = 0;
};

struct Grid {
    sm_ctx *ctx = nullptr;
    int nx2 = 0, nz2 = 0, np = 0;
    size_t N = 0, elt = 8;
    int nxh = 0, nzh = 0;           // rfft bins along x and z
    bool xpair = false;             // x derivatives by complex transforms of column pairs
    sm_dbuf spec;                   // spectrum scratch (largest of the three layouts)
    sm_dbuf kmul;                   // x multipliers (fwd, bwd), then z (fwd, bwd)
    sm_dbuf tabs;                   // per stage 2 x T_NST rows, then 2 x K_NK rows (np + 1 each)
    size_t kx_n = 0;                // entries per x multiplier row (nx2 paired, nxh otherwise)
    hipfftHandle px_f = 0, px_i = 0, pz_f = 0, pz_i = 0;
    // hand-written transforms per axis (0 x, 1 z): radices packed 4 bits a
    // pass (index into FFT_RADICES), pairs per workgroup, twiddles
    // e^{-2 pi i k/n} and whole-spectrum multipliers (fwd then bwd)
    bool own[2] = {false, false};
    int nr[2] = {0, 0};
    uint64_t rcode[2] = {0, 0};
    int P[2] = {1, 1};
    sm_dbuf tw[2], fmul[2];
};

// one layer set as the stage kernels see it
template <typename T>
struct Layer {
    T *m0, *m1, *m2;
    const T *b1, *c1, *k1;          // rows nl < np (first side)
    const T *b2, *c2, *k2;          // rows nl <= lastn (far side)
    int mode, lastn;
    unsigned lim2, lim0;
};

// the derivative value d at cell (i, j) corrected in the layer slots of L
template <typename T>
__device__ __forceinline__ T layer_fix(const Layer<T> &L, T d, int i, int j, int nx2, int nz2, int np)
{
    int nl, col, cols;
    bool first;
    if (L.mode == 0) {
        cols = nz2; col = j;
        if (i < np) { first = true; nl = i; }
        else if (i >= nx2 - 1 - L.lastn) { first = false; nl = nx2 - 1 - i; }
        else return d;
    } else {
        cols = nx2; col = i;
        if (j < np) { first = true; nl = j; }
        else if (j >= nz2 - 1 - L.lastn) { first = false; nl = nz2 - 1 - j; }
        else return d;
    }
    const unsigned q = (unsigned)((first ? 0 : np) + nl) * (unsigned)cols + (unsigned)col;
    const T b = first ? L.b1[nl] : L.b2[nl], c = first ? L.c1[nl] : L.c2[nl];
    const T k = first ? L.k1[nl] : L.k2[nl];
    const T m = b * L.m0[q] + c * d;
    L.m1[q] = m;
    const T out = d / k + L.m2[q];
    if (q < L.lim2) L.m2[q] = m;
    if (q < L.lim0) L.m0[q] = m;
    return out;
}

// complex multiply by the multiplier of bin l % m (ROWS false) or l / m
template <typename T, bool ROWS>
__global__ void pstd_kmul(Cx<T> *o, const Cx<T> *k, unsigned m, size_t n)
{
    const size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const Cx<T> a = o[l], b = k[ROWS ? l / m : l % m];
    o[l] = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

inline unsigned blocks(size_t n, unsigned bs = 256) { return (unsigned)((n + bs - 1) / bs); }

inline hipfftResult exec_r2c(hipfftHandle p, const double *in, void *out)
{
    return hipfftExecD2Z(p, const_cast<double *>(in), (hipfftDoubleComplex *)out);
}
inline hipfftResult exec_r2c(hipfftHandle p, const float *in, void *out)
{
    return hipfftExecR2C(p, const_cast<float *>(in), (hipfftComplex *)out);
}
inline hipfftResult exec_c2r(hipfftHandle p, void *in, double *out)
{
    return hipfftExecZ2D(p, (hipfftDoubleComplex *)in, out);
}
inline hipfftResult exec_c2r(hipfftHandle p, void *in, float *out)
{
    return hipfftExecC2R(p, (hipfftComplex *)in, out);
}
inline hipfftResult exec_c2c(hipfftHandle p, const double *in, void *out, int dir)
{
    return hipfftExecZ2Z(p, (hipfftDoubleComplex *)const_cast<double *>(in), (hipfftDoubleComplex *)out, dir);
}
inline hipfftResult exec_c2c(hipfftHandle p, const float *in, void *out, int dir)
{
    return hipfftExecC2C(p, (hipfftComplex *)const_cast<float *>(in), (hipfftComplex *)out, dir);
}

#define PSTD_FFT(call)                                                                   \
    do {                                                                                 \
        hipfftResult r_ = (call);                                                        \
        if (r_ != HIPFFT_SUCCESS) return fail(SM_ERR_HIP, "%s: hipFFT error %d", #call, (int)r_); \
    } while (0)

// derivative of f along axis (0 x, 1 z), forward (D+) or backward (D-),
// into out (nx2 x nz2, may not alias f)
template <typename T>
int deriv(Grid &g, const T *f, int axis, bool fwd, T *out)
{
    hipStream_t st = g.ctx->stream;
    Cx<T> *spec = (Cx<T> *)g.spec.ptr;
    const Cx<T> *km = (const Cx<T> *)g.kmul.ptr;
    if (axis == 0 && g.xpair) {
        const size_t n = g.N / 2;
        PSTD_FFT(exec_c2c(g.px_f, f, spec, HIPFFT_FORWARD));
        pstd_kmul<T, true><<<blocks(n), 256, 0, st>>>(spec, km + (fwd ? 0 : g.kx_n), (unsigned)(g.nz2 / 2), n);
        PSTD_FFT(exec_c2c(g.px_f, (const T *)spec, out, HIPFFT_BACKWARD));
    } else if (axis == 0) {
        const size_t n = (size_t)g.nxh * g.nz2;
        PSTD_FFT(exec_r2c(g.px_f, f, spec));
        pstd_kmul<T, false><<<blocks(n), 256, 0, st>>>(spec, km + (fwd ? 0 : g.kx_n), (unsigned)g.nxh, n);
        PSTD_FFT(exec_c2r(g.px_i, spec, out));
    } else {
        const size_t n = (size_t)g.nzh * g.nx2;
        PSTD_FFT(exec_r2c(g.pz_f, f, spec));
        pstd_kmul<T, false><<<blocks(n), 256, 0, st>>>(spec, km + 2 * g.kx_n + (fwd ? 0 : g.nzh),
                                                       (unsigned)g.nzh, n);
        PSTD_FFT(exec_c2r(g.pz_i, spec, out));
    }
    return SM_OK;
}

// ------------------------------------------------------------------
// hand-written FFT derivatives
// ------------------------------------------------------------------
constexpr int FFT_MAXPASS = 12;
constexpr int FFT_RADICES[9] = {2, 3, 4, 5, 7, 8, 11, 13, 16};

// radices of n, largest powers of two first; false if n has another factor
inline bool fft_factor(int n, int &nr, uint64_t &code)
{
    nr = 0;
    code = 0;
    int m = n;
    auto push = [&](int r) {
        int q = 0;
        while (FFT_RADICES[q] != r) ++q;
        code |= (uint64_t)q << (4 * nr);
        ++nr;
    };
    // odd radices first: the first Stockham pass writes with a stride of R
    // elements, which for a power of two puts a wave's lanes on a few LDS
    // banks; later passes write runs of Ns consecutive elements
    int two = 0;
    while (m % 2 == 0) { m /= 2; ++two; }
    for (int r : {13, 11, 7, 5, 3})
        while (m % r == 0) {
            if (nr >= FFT_MAXPASS) return false;
            push(r);
            m /= r;
        }
    if (two % 4 == 3) push(8);
    else if (two % 4 == 2) push(4);
    else if (two % 4 == 1) push(2);
    for (int q = 0; q < two / 4; ++q) {
        if (nr >= FFT_MAXPASS) return false;
        push(16);
    }
    return m == 1 && n >= 2 && nr > 0 && nr <= FFT_MAXPASS;
}

// a b + c with one rounding
__device__ __forceinline__ float mad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double mad(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename T>
__device__ __forceinline__ Cx<T> cmul(Cx<T> a, Cx<T> b)
{
    // fused multiply-adds: one rounding less per product (the transforms
    // are held to a tolerance, not to the reference's FFTW bits)
    return {mad(a.re, b.re, -(a.im * b.im)), mad(a.re, b.im, a.im * b.re)};
}
template <typename T>
__device__ __forceinline__ Cx<T> cadd(Cx<T> a, Cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ __forceinline__ Cx<T> csub(Cx<T> a, Cx<T> b) { return {a.re - b.re, a.im - b.im}; }
// -i a
template <typename T>
__device__ __forceinline__ Cx<T> cmi(Cx<T> a) { return {a.im, -a.re}; }

// forward DFT of v[0..R) in place, w = e^{-2 pi i / R} = r[rs] (r holds
// the roots of unity of the pass's radix; rs > 1 inside a larger radix)
template <typename T, int R>
__device__ __forceinline__ void dft(Cx<T> *v, const Cx<T> *r, int rs)
{
    if constexpr (R == 2) {
        const Cx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b); v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        const Cx<T> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]), a2 = cadd(v[1], v[3]), a3 = csub(v[1], v[3]);
        v[0] = cadd(a0, a2); v[2] = csub(a0, a2);
        v[1] = cadd(a1, cmi(a3)); v[3] = csub(a1, cmi(a3));
    } else if constexpr (R == 8 || R == 16) {
        // two half-length DFTs of the even and odd samples, combined with
        // w^k (from the table)
        constexpr int H = R / 2;
        Cx<T> e[H], o[H];
#pragma unroll
        for (int t = 0; t < H; ++t) { e[t] = v[2 * t]; o[t] = v[2 * t + 1]; }
        dft<T, H>(e, r, 2 * rs);
        dft<T, H>(o, r, 2 * rs);
#pragma unroll
        for (int k = 0; k < H; ++k) {
            const Cx<T> wo = k == 0 ? o[k] : cmul(o[k], r[k * rs]);
            v[k] = cadd(e[k], wo);
            v[k + H] = csub(e[k], wo);
        }
    } else {
        // odd R: pair t with R - t; X_s = A_s - i B_s, X_{R-s} = A_s + i B_s
        // with A_s = v0 + sum (v_t + v_{R-t}) cos, B_s = sum (v_t - v_{R-t}) sin
        constexpr int H = (R - 1) / 2;
        Cx<T> sp[H], sm[H];
        Cx<T> x0 = v[0];
#pragma unroll
        for (int t = 1; t <= H; ++t) {
            sp[t - 1] = cadd(v[t], v[R - t]);
            sm[t - 1] = csub(v[t], v[R - t]);
            x0 = cadd(x0, sp[t - 1]);
        }
        Cx<T> out[R];
        out[0] = x0;
#pragma unroll
        for (int s = 1; s <= H; ++s) {
            Cx<T> A = v[0], B = {T(0), T(0)};
#pragma unroll
            for (int t = 1; t <= H; ++t) {
                const Cx<T> w = r[((s * t) % R) * rs];   // (cos, -sin)
                A.re = mad(sp[t - 1].re, w.re, A.re); A.im = mad(sp[t - 1].im, w.re, A.im);
                B.re = mad(-sm[t - 1].re, w.im, B.re); B.im = mad(-sm[t - 1].im, w.im, B.im);
            }
            out[s] = {A.re + B.im, A.im - B.re};
            out[R - s] = {A.re - B.im, A.im + B.re};
        }
#pragma unroll
        for (int t = 0; t < R; ++t) v[t] = out[t];
    }
}

// where a pass reads and writes: the LDS buffers (transform p at p ls), or,
// for the first forward and the last inverse pass, the field f and the
// derivative out in global memory (pair p: rows or columns r0, r0 + 1 with
// r0 = r0base + 2p, a missing partner read as 0 and not written); with mul,
// the pass writes conj(mul[k] v) (the multiplier between the transforms)
template <typename T>
struct PassIO {
    const Cx<T> *src;
    Cx<T> *dst;
    const T *f;
    T *out;
    const Cx<T> *mul;
    int ls, r0base, lim, nz2;
    bool in_g, out_g;
};

template <typename T, int AX>
__device__ __forceinline__ Cx<T> gload(const PassIO<T> io, int p, int k)
{
    const int r0 = io.r0base + 2 * p;
    T re = T(0), im = T(0);
    if (AX == 1) {
        if (r0 < io.lim) re = io.f[(size_t)r0 * io.nz2 + k];
        if (r0 + 1 < io.lim) im = io.f[(size_t)(r0 + 1) * io.nz2 + k];
    } else {
        if (r0 < io.lim) re = io.f[(size_t)k * io.nz2 + r0];
        if (r0 + 1 < io.lim) im = io.f[(size_t)k * io.nz2 + r0 + 1];
    }
    return {re, im};
}

template <typename T, int AX>
__device__ __forceinline__ void gstore(const PassIO<T> io, int p, int k, Cx<T> y)
{
    const int r0 = io.r0base + 2 * p;
    if (AX == 1) {
        if (r0 < io.lim) io.out[(size_t)r0 * io.nz2 + k] = y.re;
        if (r0 + 1 < io.lim) io.out[(size_t)(r0 + 1) * io.nz2 + k] = -y.im;
    } else {
        if (r0 < io.lim) io.out[(size_t)k * io.nz2 + r0] = y.re;
        if (r0 + 1 < io.lim) io.out[(size_t)k * io.nz2 + r0 + 1] = -y.im;
    }
}

// one Stockham pass of radix R over P transforms of length n; Ns = product
// of the earlier radices; twp[(t - 1) Ns + k] = e^{-2 pi i k t / (Ns R)}
// (consecutive k on consecutive LDS banks), roots the R-th roots of unity.
// Passes that touch global memory along x take the pairs fastest (2P
// adjacent values of a row per lane group).
template <typename T, int R, int AX>
__device__ __forceinline__ void fft_pass(const PassIO<T> io, int n, int P, int Ns, const Cx<T> *twp,
                                         const Cx<T> *roots)
{
    const int nb = n / R, ls = io.ls;
    const bool pf = AX == 0 && (io.in_g || io.out_g);
    for (int l = threadIdx.x; l < nb * P; l += blockDim.x) {
        int p, j;
        if (pf) { j = l / P; p = l - j * P; }
        else { p = l / nb; j = l - p * nb; }
        const int k = j % Ns;
        Cx<T> v[R];
        if (io.in_g) {
#pragma unroll
            for (int t = 0; t < R; ++t) v[t] = gload<T, AX>(io, p, j + t * nb);
        } else {
            const Cx<T> *s = io.src + p * ls;
#pragma unroll
            for (int t = 0; t < R; ++t) v[t] = s[j + t * nb];
        }
        if (k) {
#pragma unroll
            for (int t = 1; t < R; ++t) v[t] = cmul(v[t], twp[(t - 1) * Ns + k]);
        }
        dft<T, R>(v, roots, 1);
        const int base = (j - k) * R + k;
        if (io.mul) {
#pragma unroll
            for (int t = 0; t < R; ++t) {
                const Cx<T> z = cmul(v[t], io.mul[base + t * Ns]);
                v[t] = {z.re, -z.im};
            }
        }
        if (io.out_g) {
#pragma unroll
            for (int t = 0; t < R; ++t) gstore<T, AX>(io, p, base + t * Ns, v[t]);
        } else {
            Cx<T> *d = io.dst + p * ls;
#pragma unroll
            for (int t = 0; t < R; ++t) d[base + t * Ns] = v[t];
        }
    }
}

// a forward FFT of P transforms: from global memory (in_g) or LDS buffer a
// (scratch b), into LDS (returned, after a barrier) or global memory
// (out_g), the last pass applying mul if given.  Inlined: an out-of-line
// call would keep a stack frame in scratch memory.
template <typename T, int AX>
__device__ __forceinline__ Cx<T> *fft_run(PassIO<T> io, Cx<T> *a, Cx<T> *b, int n, int P, int nr, uint64_t code,
                                          const Cx<T> *tw, bool in_g, bool out_g, const Cx<T> *mul)
{
    int Ns = 1;
    for (int q = 0; q < nr; ++q) {
        const int ri = (int)((code >> (4 * q)) & 15);
        const int R = FFT_RADICES[ri];
        const Cx<T> *roots = tw + (R - 1) * Ns;
        io.src = a;
        io.dst = b;
        io.in_g = in_g && q == 0;
        io.out_g = out_g && q == nr - 1;
        io.mul = q == nr - 1 ? mul : nullptr;
        switch (ri) {
        case 0: fft_pass<T, 2, AX>(io, n, P, Ns, tw, roots); break;
        case 1: fft_pass<T, 3, AX>(io, n, P, Ns, tw, roots); break;
        case 2: fft_pass<T, 4, AX>(io, n, P, Ns, tw, roots); break;
        case 3: fft_pass<T, 5, AX>(io, n, P, Ns, tw, roots); break;
        case 4: fft_pass<T, 7, AX>(io, n, P, Ns, tw, roots); break;
        case 5: fft_pass<T, 8, AX>(io, n, P, Ns, tw, roots); break;
        case 6: fft_pass<T, 11, AX>(io, n, P, Ns, tw, roots); break;
        case 7: fft_pass<T, 13, AX>(io, n, P, Ns, tw, roots); break;
        default: fft_pass<T, 16, AX>(io, n, P, Ns, tw, roots); break;
        }
        tw = roots + R;   // the next pass's table
        Ns *= R;
        if (!io.out_g) __syncthreads();
        Cx<T> *t = a; a = b; b = t;
    }
    return a;
}

// up to 8 derivatives of one stage along one axis: job j differentiates
// field ((fields >> 4j) & 15) of x into d + ((slots >> 4j) & 15) N, with
// the forward multiplier if bit j of fwd is set
template <typename T>
struct DfftArgs {
    const T *x;
    T *d;
    unsigned N;
    uint32_t fields, slots, fwd;
    int njobs;
    const Cx<T> *tw, *mul_f, *mul_b;
    int n, nr, tsz;
    uint64_t code;
    int nx2, nz2;
    int P, ntiles, nblocks;
};

extern __shared__ __align__(16) unsigned char pstd_lds[];

// AX 1: z (unit stride), row pairs (r, r + 1); AX 0: x (stride nz2),
// column pairs (c, c + 1); a missing partner reads as 0 and is not written.
// b: the workgroup's tile of this axis's jobs
template <typename T, int AX>
__device__ __forceinline__ void dfft_body(const DfftArgs<T> &a, int b)
{
    const int jb = b / a.ntiles, tile = b - jb * a.ntiles;
    const T *f = a.x + (size_t)((a.fields >> (4 * jb)) & 15) * a.N;
    T *out = a.d + (size_t)((a.slots >> (4 * jb)) & 15) * a.N;
    const Cx<T> *mul = ((a.fwd >> jb) & 1) ? a.mul_f : a.mul_b;
    const int n = a.n, P = a.P, ls = n + 1;
    // LDS: two buffers of P transforms (stride n + 1), then the per-pass
    // twiddles and roots
    Cx<T> *A = (Cx<T> *)pstd_lds, *B = A + P * ls, *tw = B + P * ls;
    for (int l = threadIdx.x; l < a.tsz; l += blockDim.x) tw[l] = a.tw[l];
    __syncthreads();
    PassIO<T> io;
    io.f = f; io.out = out; io.ls = ls; io.r0base = 2 * tile * P;
    io.lim = AX == 1 ? a.nx2 : a.nz2; io.nz2 = a.nz2;
    // Z = DFT(f) from global memory, written as conj(M Z); y = IDFT(M Z) =
    // conj(DFT(conj(M Z))) into global memory (one call site: inlined once)
    Cx<T> *c = A;
    for (int rep = 0; rep < 2; ++rep)
        c = fft_run<T, AX>(io, c, c == A ? B : A, n, P, a.nr, a.code, tw, rep == 0, rep == 1,
                           rep == 0 ? mul : nullptr);
}

// workgroups of one XCD take consecutive tiles (adjacent columns share
// cache lines)
__device__ __forceinline__ int xcd_tile(int nblocks)
{
    const int per = (nblocks + 7) / 8;
    return (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
}

template <typename T, int AX>
__global__ __launch_bounds__(512) void pstd_dfft(const DfftArgs<T> a)
{
    const int b = xcd_tile(a.nblocks);
    if (b < a.nblocks) dfft_body<T, AX>(a, b);
}

// the derivatives of a stage along x (ax) and z (az) in one grid: the x
// tiles first, then the z tiles
template <typename T>
__global__ __launch_bounds__(512) void pstd_dfft_xz(const DfftArgs<T> ax, const DfftArgs<T> az)
{
    const int b = xcd_tile(ax.nblocks + az.nblocks);
    if (b < ax.nblocks) dfft_body<T, 0>(ax, b);
    else if (b < ax.nblocks + az.nblocks) dfft_body<T, 1>(az, b - ax.nblocks);
}

// sources the stage kernels inject themselves, after the update of their
// cell (at most SRC_FUSED: more go through the solver's separate
// one-thread source kernel): cell[s] of source s; type, cell, offset and
// length of all nsrc sources in the device table tab, the functions in fct
constexpr int SRC_FUSED = 8;
template <typename T>
struct StageSrc {
    int n;                       // fused sources (0: none, or injected separately)
    int cell[SRC_FUSED];
    const int64_t *tab;
    const T *fct;
    int nsrc;
    T dtw, dtt;                  // the stage's weights of the source function
    int64_t it;
};

struct DJob {
    int field, slot;
    bool fwd;
};

// complex entries of the per-pass tables: (R - 1) Ns twiddles and R roots
inline int tw_count(int nr, uint64_t code)
{
    int c = 0, Ns = 1;
    for (int q = 0; q < nr; ++q) {
        const int R = FFT_RADICES[(code >> (4 * q)) & 15];
        c += (R - 1) * Ns + R;
        Ns *= R;
    }
    return c;
}

inline size_t dfft_lds(int n, int P, size_t elt, int tsz)
{
    return (2 * (size_t)P * (n + 1) + (size_t)tsz) * 2 * elt;
}

// threads of a pstd_dfft workgroup: enough for the largest pass in one
// round (P n / R butterflies), at most 512
inline unsigned dfft_threads(int n, int P, int nr, uint64_t code)
{
    int most = 0;
    for (int q = 0; q < nr; ++q) most = std::max(most, P * n / FFT_RADICES[(code >> (4 * q)) & 15]);
    return (unsigned)std::min(512, std::max(64, (most + 63) / 64 * 64));
}

// the arguments, LDS bytes and threads of the derivatives jobs[0..nj)
// along axis of the fields of x into d (own transforms)
template <typename T>
DfftArgs<T> dfft_args(const Grid &g, int axis, const T *x, T *d, const DJob *jobs, int nj, size_t &lds,
                      unsigned &nt)
{
    DfftArgs<T> a;
    a.x = x; a.d = d; a.N = (unsigned)g.N;
    a.fields = a.slots = a.fwd = 0;
    for (int q = 0; q < nj; ++q) {
        a.fields |= (uint32_t)jobs[q].field << (4 * q);
        a.slots |= (uint32_t)jobs[q].slot << (4 * q);
        a.fwd |= (uint32_t)(jobs[q].fwd ? 1 : 0) << q;
    }
    a.njobs = nj;
    const int n = axis == 0 ? g.nx2 : g.nz2, other = axis == 0 ? g.nz2 : g.nx2;
    a.tw = (const Cx<T> *)g.tw[axis].ptr;
    a.mul_f = (const Cx<T> *)g.fmul[axis].ptr;
    a.mul_b = a.mul_f + n;
    a.n = n; a.nr = g.nr[axis]; a.code = g.rcode[axis]; a.tsz = tw_count(a.nr, a.code);
    a.nx2 = g.nx2; a.nz2 = g.nz2;
    a.P = g.P[axis];
    a.ntiles = ((other + 1) / 2 + a.P - 1) / a.P;
    a.nblocks = nj * a.ntiles;
    lds = dfft_lds(n, a.P, sizeof(T), a.tsz);
    nt = dfft_threads(n, a.P, a.nr, a.code);
    return a;
}

// the derivatives jobs[0..nj) along axis of the fields of x into d
template <typename T>
int deriv_batch(Grid &g, int axis, const T *x, T *d, const DJob *jobs, int nj)
{
    if (!g.own[axis]) {
        for (int q = 0; q < nj; ++q) {
            int rc = deriv<T>(g, x + (size_t)jobs[q].field * g.N, axis, jobs[q].fwd, d + (size_t)jobs[q].slot * g.N);
            if (rc != SM_OK) return rc;
        }
        return SM_OK;
    }
    size_t lds;
    unsigned nt;
    const DfftArgs<T> a = dfft_args<T>(g, axis, x, d, jobs, nj, lds, nt);
    const unsigned grid = (unsigned)((a.nblocks + 7) / 8 * 8);
    if (axis == 0) pstd_dfft<T, 0><<<grid, nt, lds, g.ctx->stream>>>(a);
    else pstd_dfft<T, 1><<<grid, nt, lds, g.ctx->stream>>>(a);
    return SM_OK;
}

// the derivatives xj[0..nx) along x and zj[0..nz) along z of the fields
// of x into d: one launch when both axes use the own transforms
template <typename T>
int deriv_xz(Grid &g, const T *x, T *d, const DJob *xj, int nx, const DJob *zj, int nz)
{
    if (!(g.own[0] && g.own[1] && nx > 0 && nz > 0)) {
        int rc = nx ? deriv_batch<T>(g, 0, x, d, xj, nx) : SM_OK;
        if (rc == SM_OK && nz) rc = deriv_batch<T>(g, 1, x, d, zj, nz);
        return rc;
    }
    size_t lx, lz;
    unsigned tx, tz;
    const DfftArgs<T> ax = dfft_args<T>(g, 0, x, d, xj, nx, lx, tx);
    const DfftArgs<T> az = dfft_args<T>(g, 1, x, d, zj, nz, lz, tz);
    const unsigned grid = (unsigned)((ax.nblocks + az.nblocks + 7) / 8 * 8);
    pstd_dfft_xz<T><<<grid, std::max(tx, tz), std::max(lx, lz), g.ctx->stream>>>(ax, az);
    return SM_OK;
}

// device views of the CPML tables: stage 0 (fp1), 1 (fp23), 2 (fp4); axis 0 x, 1 z
template <typename T>
const T *tab(const Grid &g, int stage, int axis, int row)
{
    return (const T *)g.tabs.ptr + ((size_t)(stage * 2 + axis) * T_NST + row) * (g.np + 1);
}
template <typename T>
const T *ktab(const Grid &g, int axis, int row)
{
    return (const T *)g.tabs.ptr + ((size_t)3 * 2 * T_NST + axis * K_NK + row) * (g.np + 1);
}

// a layer set: mode, last far-side row, tables (stage, axis, rows) of the
// first and far sides, memory copies, copy limits
template <typename T>
Layer<T> layer(const Grid &g, int stage, T *m0, T *m1, T *m2, int mode, int lastn, int ax1, int b1, int kk1,
               int ax2, int b2, int kk2, size_t lim2, size_t lim0)
{
    Layer<T> L;
    L.m0 = m0; L.m1 = m1; L.m2 = m2;
    L.b1 = tab<T>(g, stage, ax1, b1); L.c1 = tab<T>(g, stage, ax1, b1 + 1); L.k1 = ktab<T>(g, ax1, kk1);
    L.b2 = tab<T>(g, stage, ax2, b2); L.c2 = tab<T>(g, stage, ax2, b2 + 1); L.k2 = ktab<T>(g, ax2, kk2);
    L.mode = mode; L.lastn = lastn;
    L.lim2 = (unsigned)lim2; L.lim0 = (unsigned)lim0;
    return L;
}

// CPML tables of pml.c:1959-2106 for one axis
inline void cpml_axis(const Cpml &p, size_t n, double d, double dt, double alpha_max,
                      std::vector<double> st[3][T_NST], std::vector<double> kk[K_NK])
{
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
            const double h = s == 1 ? dt * 0.5 : dt;
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

inline int upload_any(size_t elt, void *dst, const std::vector<double> &v, hipStream_t st)
{
    return elt == 8 ? upload<double>(dst, v, st) : upload<float>(dst, v, st);
}

inline int make_plan(hipfftHandle *h, int n, int howmany, int istride, int idist, int ostride, int odist,
                     int in_total, int out_total, hipfftType type, hipStream_t st)
{
    int nn[1] = {n}, inembed[1] = {in_total}, onembed[1] = {out_total};
    hipfftResult r = hipfftPlanMany(h, 1, nn, inembed, istride, idist, onembed, ostride, odist, type, howmany);
    if (r != HIPFFT_SUCCESS) return fail(SM_ERR_HIP, "hipfftPlanMany (n %d, batch %d): error %d", n, howmany, (int)r);
    r = hipfftSetStream(*h, st);
    if (r != HIPFFT_SUCCESS) return fail(SM_ERR_HIP, "hipfftSetStream: error %d", (int)r);
    return SM_OK;
}

inline void release(Grid &g)
{
    for (hipfftHandle *h : {&g.px_f, &g.px_i, &g.pz_f, &g.pz_i})
        if (*h) { (void)hipfftDestroy(*h); *h = 0; }
    for (sm_dbuf *b : {&g.spec, &g.kmul, &g.tabs, &g.tw[0], &g.tw[1], &g.fmul[0], &g.fmul[1]}) dfree(*b);
}

inline size_t spec_bytes(size_t nx2, size_t nz2, size_t elt)
{
    const size_t a = (nx2 / 2 + 1) * nz2, b = nx2 * (nz2 / 2 + 1);
    return std::max(a, b) * 2 * elt;
}

inline size_t table_bytes(size_t np, size_t elt) { return (3 * 2 * T_NST + 2 * K_NK) * (np + 1) * elt; }

inline size_t kmul_bytes(size_t nx2, size_t nz2, size_t elt)
{
    return 2 * (std::max(nx2, nx2 / 2 + 1) + nz2 / 2 + 1) * 2 * elt;
}

// plans, multipliers and CPML tables for an nx2 x nz2 grid with np layers;
// own: hand-written transforms on the axes whose length allows them
inline int init(Grid &g, sm_ctx *ctx, size_t nx2, size_t nz2, size_t np, size_t elt, double dx, double dz,
                double dt, double f0, const Cpml &cp, bool own = true)
{
    g.ctx = ctx;
    g.nx2 = (int)nx2; g.nz2 = (int)nz2; g.np = (int)np;
    g.N = nx2 * nz2; g.elt = elt;
    g.nxh = (int)(nx2 / 2 + 1); g.nzh = (int)(nz2 / 2 + 1);
    g.xpair = nz2 % 2 == 0;
    g.kx_n = g.xpair ? nx2 : (size_t)g.nxh;
    hipStream_t st = ctx->stream;
    int rc = dalloc(g.spec, spec_bytes(nx2, nz2, elt));
    if (rc == SM_OK) rc = dalloc(g.kmul, kmul_bytes(nx2, nz2, elt));
    if (rc == SM_OK) rc = dalloc(g.tabs, table_bytes(np, elt));
    if (rc != SM_OK) return rc;
    // multipliers i k e^{+-i k d/2}, k = 2 pi n / (d n2^2) (propagate.c:363-374)
    {
        std::vector<double> km;
        auto h = [](size_t n2, double d, double sgn, size_t n, double &re, double &im) {
            const double dn = 8.0 * std::atan(1.0) / (d * n2 * n2);
            const double k = n * dn, ph = sgn * 0.5 * k * d * n2;
            re = -k * std::sin(ph);   // re(i k e^{i ph})
            im = k * std::cos(ph);
        };
        auto add_half = [&](size_t n2, double d, double sgn) {
            for (size_t n = 0; n < n2 / 2 + 1; ++n) {
                double re, im;
                h(n2, d, sgn, n, re, im);
                km.push_back(re); km.push_back(im);
            }
        };
        auto add_full = [&](size_t n2, double d, double sgn) {
            for (size_t n = 0; n < n2; ++n) {
                double re, im;
                if (2 * n < n2) h(n2, d, sgn, n, re, im);
                else if (2 * n == n2) { h(n2, d, sgn, n, re, im); im = 0.0; }
                else { h(n2, d, sgn, n2 - n, re, im); im = -im; }
                km.push_back(re); km.push_back(im);
            }
        };
        if (g.xpair) { add_full(nx2, dx, 1.0); add_full(nx2, dx, -1.0); }
        else { add_half(nx2, dx, 1.0); add_half(nx2, dx, -1.0); }
        add_half(nz2, dz, 1.0); add_half(nz2, dz, -1.0);
        rc = upload_any(elt, g.kmul.ptr, km, st);
    }
    // CPML tables
    if (rc == SM_OK) {
        const double pi = 4.0 * std::atan(1.0);
        const double alpha_max = cp.alpha_max ? pi * f0 : 0.0;
        std::vector<double> sx[3][T_NST], sz[3][T_NST], kx[K_NK], kz[K_NK];
        cpml_axis(cp, np, dx, dt, alpha_max, sx, kx);
        cpml_axis(cp, np, dz, dt, alpha_max, sz, kz);
        std::vector<double> t;
        for (int s3 = 0; s3 < 3; ++s3) {
            for (int r = 0; r < T_NST; ++r) t.insert(t.end(), sx[s3][r].begin(), sx[s3][r].end());
            for (int r = 0; r < T_NST; ++r) t.insert(t.end(), sz[s3][r].begin(), sz[s3][r].end());
        }
        for (int r = 0; r < K_NK; ++r) t.insert(t.end(), kx[r].begin(), kx[r].end());
        for (int r = 0; r < K_NK; ++r) t.insert(t.end(), kz[r].begin(), kz[r].end());
        rc = upload_any(elt, g.tabs.ptr, t, st);
    }
    // hand-written transforms: twiddles e^{-2 pi i k/n} and the multipliers
    // over the whole spectrum; P pairs per workgroup within 80 KiB of LDS
    // (two workgroups per CU)
    int lds_max = 65536;
    if (own) {
        const size_t want = 81920;
        const bool ok = elt == 8
            ? hipFuncSetAttribute((const void *)pstd_dfft<double, 0>, hipFuncAttributeMaxDynamicSharedMemorySize, want) == hipSuccess &&
              hipFuncSetAttribute((const void *)pstd_dfft<double, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, want) == hipSuccess &&
              hipFuncSetAttribute((const void *)pstd_dfft_xz<double>, hipFuncAttributeMaxDynamicSharedMemorySize, want) == hipSuccess
            : hipFuncSetAttribute((const void *)pstd_dfft<float, 0>, hipFuncAttributeMaxDynamicSharedMemorySize, want) == hipSuccess &&
              hipFuncSetAttribute((const void *)pstd_dfft<float, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, want) == hipSuccess &&
              hipFuncSetAttribute((const void *)pstd_dfft_xz<float>, hipFuncAttributeMaxDynamicSharedMemorySize, want) == hipSuccess;
        (void)hipGetLastError();
        if (ok) lds_max = (int)want;
    }
    for (int ax = 0; ax < 2 && rc == SM_OK && own; ++ax) {
        const size_t n = ax == 0 ? nx2 : nz2;
        const double d = ax == 0 ? dx : dz;
        if (!fft_factor((int)n, g.nr[ax], g.rcode[ax])) continue;
        const int tsz = tw_count(g.nr[ax], g.rcode[ax]);
        if (dfft_lds((int)n, 1, elt, tsz) > (size_t)lds_max) continue;
        // 4 column or row pairs per workgroup where LDS allows (adjacent
        // x columns share cache lines; measured on 1040-point axes, FP32:
        // 4/4 pairs 9% faster per step than 4/2, 2/2 and 1/1 slower still)
        g.P[ax] = std::max(1, std::min(4, (int)(lds_max / dfft_lds((int)n, 1, elt, tsz))));
        while (g.P[ax] > 1 && dfft_lds((int)n, g.P[ax], elt, tsz) > (size_t)lds_max) --g.P[ax];
        std::vector<double> tw, fm;
        const double two_pi = 8.0 * std::atan(1.0);
        // per pass: e^{-2 pi i k t / (Ns R)} at (t - 1) Ns + k, then the
        // R-th roots of unity
        for (int q = 0, Ns = 1; q < g.nr[ax]; ++q) {
            const int R = FFT_RADICES[(g.rcode[ax] >> (4 * q)) & 15];
            for (int t = 1; t < R; ++t)
                for (int k = 0; k < Ns; ++k) {
                    const double th = two_pi * (double)k * t / ((double)Ns * R);
                    tw.push_back(std::cos(th)); tw.push_back(-std::sin(th));
                }
            for (int m = 0; m < R; ++m) {
                const double th = two_pi * (double)m / R;
                tw.push_back(std::cos(th)); tw.push_back(-std::sin(th));
            }
            Ns *= R;
        }
        for (double sgn : {1.0, -1.0}) {
            const double dn = two_pi / (d * n * n);
            auto h = [&](size_t m, double &re, double &im) {
                const double k = m * dn, ph = sgn * 0.5 * k * d * n;
                re = -k * std::sin(ph); im = k * std::cos(ph);
            };
            for (size_t m = 0; m < n; ++m) {
                double re, im;
                if (2 * m < n) h(m, re, im);
                else if (2 * m == n) { h(m, re, im); im = 0.0; }
                else { h(n - m, re, im); im = -im; }
                fm.push_back(re); fm.push_back(im);
            }
        }
        rc = dalloc(g.tw[ax], tw.size() * elt);
        if (rc == SM_OK) rc = dalloc(g.fmul[ax], fm.size() * elt);
        if (rc == SM_OK) rc = upload_any(elt, g.tw[ax].ptr, tw, st);
        if (rc == SM_OK) rc = upload_any(elt, g.fmul[ax].ptr, fm, st);
        if (rc == SM_OK) g.own[ax] = true;
    }
    // plans (propagate.c:386-497) for the other axes: x transforms of stride
    // nz2 (paired: complex, stride nz2/2), z transforms of unit stride
    if (rc == SM_OK) {
        const hipfftType f_t = elt == 8 ? HIPFFT_D2Z : HIPFFT_R2C, i_t = elt == 8 ? HIPFFT_Z2D : HIPFFT_C2R;
        const hipfftType c_t = elt == 8 ? HIPFFT_Z2Z : HIPFFT_C2C;
        const int NX = (int)nx2, NZ = (int)nz2, NN = (int)g.N;
        if (g.own[0]) {
        } else if (g.xpair) {
            rc = make_plan(&g.px_f, NX, NZ / 2, NZ / 2, 1, NZ / 2, 1, NN / 2, NN / 2, c_t, st);
        } else {
            rc = make_plan(&g.px_f, NX, NZ, NZ, 1, 1, g.nxh, NN, g.nxh * NZ, f_t, st);
            if (rc == SM_OK) rc = make_plan(&g.px_i, NX, NZ, 1, g.nxh, NZ, 1, g.nxh * NZ, NN, i_t, st);
        }
        if (rc == SM_OK && !g.own[1]) {
            rc = make_plan(&g.pz_f, NZ, NX, 1, NZ, 1, g.nzh, NN, NX * g.nzh, f_t, st);
            if (rc == SM_OK) rc = make_plan(&g.pz_i, NZ, NX, 1, g.nzh, 1, NZ, NX * g.nzh, NN, i_t, st);
        }
    }
    return rc;
}

// device bytes of the machinery
inline size_t grid_bytes(size_t nx2, size_t nz2, size_t np, size_t elt)
{
    // per axis: twiddles and roots (n - 1 + the sum of the radices) and two
    // multiplier tables of n
    return spec_bytes(nx2, nz2, elt) + kmul_bytes(nx2, nz2, elt) + table_bytes(np, elt) +
           (3 * (nx2 + nz2) + 2 * 16 * FFT_MAXPASS) * 2 * elt;
}

inline size_t grid_device_bytes(const Grid &g)
{
    size_t b = g.spec.bytes + g.kmul.bytes + g.tabs.bytes;
    for (int ax = 0; ax < 2; ++ax) b += g.tw[ax].bytes + g.fmul[ax].bytes;
    return b;
}

}  // namespace pstd
}  // namespace sm
