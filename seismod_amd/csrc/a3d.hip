// a3d.hip -- 3D acoustic CPML solver (replaces a_iso_3d.c:368-417, 494-693).
//
// Host side: FP64 set-up in the reference's order (CFL, CPML tables, K*dt),
// device buffers, the time loop and the record buffers.
// Device side: the "naive" two-kernel step (velocity sweep, pressure sweep)
// in FP64 (reference operation order) and FP32, plus the fused single-pass
// step in a3d_fused.hip.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "a3d_common.hpp"
#include "a3d_fused.hpp"
#include "sm_internal.hpp"

namespace sm {

template <typename T>
struct A3dConst {
    T dx, dy, dz;      // FP64 path: spacings; FP32 path: 1/spacing
    T dt2;             // FP64: dt; FP32: 2*dt
};

// ---------------------------------------------------------------------
// naive kernels: one thread per cell, velocity then pressure
// ---------------------------------------------------------------------
template <typename T, bool EXACT>
__device__ inline void vel_component(T *v, T *mem, int s, int n, size_t c, size_t step,
                                     const T *P, const T *rho, const CpmlAxis<T> &ax,
                                     T h, T dt2)
{
    // a_iso_3d.c:611-615 (x), 628-632 (y), 644-648 (z)
    T dP;
    if constexpr (EXACT) dP = (P[c + step] - P[c]) / h;
    else dP = (P[c + step] - P[c]) * h;
    if (s >= 0) {
        T m = ax.bh[n] * mem[0] + ax.ah[n] * dP;
        mem[0] = m;
        if constexpr (EXACT) dP = dP / ax.kh[n] + m;
        else dP = dP * ax.kh[n] + m;
    }
    T irho;
    if constexpr (EXACT) irho = dt2 * 2. / (rho[c] + rho[c + step]);
    else irho = dt2 / (rho[c] + rho[c + step]);
    v[c] -= irho * dP;
}

template <typename T, bool EXACT>
__global__ void __launch_bounds__(256)
a3d_vel_naive(A3dDev<T> d, CpmlAxis<T> cx, CpmlAxis<T> cy, CpmlAxis<T> cz, A3dConst<T> k)
{
    const int kk = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int i = blockIdx.z;
    if (kk >= d.nz2 || j >= d.ny2) return;
    const size_t sx = (size_t)d.ny2 * d.nz2, sy = d.nz2;
    const size_t c = ((size_t)i * d.ny2 + j) * d.nz2 + kk;
    const int np = d.np, np2 = 2 * np;
    if (i < d.nx2 - 1) {
        int s = slab_vel(i, d.nx2, np);
        T *mem = s >= 0 ? d.mPx + ((size_t)s * d.ny2 + j) * d.nz2 + kk : nullptr;
        vel_component<T, EXACT>(d.vx, mem, s, i, c, sx, d.P, d.rho, cx, k.dx, k.dt2);
    }
    if (j < d.ny2 - 1) {
        int s = slab_vel(j, d.ny2, np);
        T *mem = s >= 0 ? d.mPy + ((size_t)i * np2 + s) * d.nz2 + kk : nullptr;
        vel_component<T, EXACT>(d.vy, mem, s, j, c, sy, d.P, d.rho, cy, k.dy, k.dt2);
    }
    if (kk < d.nz2 - 1) {
        int s = slab_vel(kk, d.nz2, np);
        T *mem = s >= 0 ? d.mPz + ((size_t)i * d.ny2 + j) * np2 + s : nullptr;
        vel_component<T, EXACT>(d.vz, mem, s, kk, c, 1, d.P, d.rho, cz, k.dz, k.dt2);
    }
}

template <typename T, bool EXACT>
__device__ inline T pre_component(T dv, T *mem, int s, int n, const CpmlAxis<T> &ax, T h)
{
    // a_iso_3d.c:667-678
    if constexpr (EXACT) dv = dv / h;
    else dv = dv * h;
    if (s >= 0) {
        T m = ax.b[n] * mem[0] + ax.a[n] * dv;
        mem[0] = m;
        if constexpr (EXACT) dv = dv / ax.k[n] + m;
        else dv = dv * ax.k[n] + m;
    }
    return dv;
}

template <typename T, bool EXACT>
__global__ void __launch_bounds__(256)
a3d_pre_naive(A3dDev<T> d, CpmlAxis<T> cx, CpmlAxis<T> cy, CpmlAxis<T> cz, A3dConst<T> k)
{
    const int kk = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int i = blockIdx.z;
    // a_iso_3d.c:659-661: i, j, k >= 1
    if (kk >= d.nz2 || j >= d.ny2 || i < 1 || j < 1 || kk < 1) return;
    const size_t sx = (size_t)d.ny2 * d.nz2, sy = d.nz2;
    const size_t c = ((size_t)i * d.ny2 + j) * d.nz2 + kk;
    const int np = d.np, np2 = 2 * np;
    int s;
    s = slab_pre(i, d.nx2, np);
    T dvdx = pre_component<T, EXACT>(d.vx[c] - d.vx[c - sx],
                                     s >= 0 ? d.mvx + ((size_t)s * d.ny2 + j) * d.nz2 + kk : nullptr,
                                     s, i, cx, k.dx);
    s = slab_pre(j, d.ny2, np);
    T dvdy = pre_component<T, EXACT>(d.vy[c] - d.vy[c - sy],
                                     s >= 0 ? d.mvy + ((size_t)i * np2 + s) * d.nz2 + kk : nullptr,
                                     s, j, cy, k.dy);
    s = slab_pre(kk, d.nz2, np);
    T dvdz = pre_component<T, EXACT>(d.vz[c] - d.vz[c - 1],
                                     s >= 0 ? d.mvz + ((size_t)i * d.ny2 + j) * np2 + s : nullptr,
                                     s, kk, cz, k.dz);
    d.P[c] -= d.kdt[c] * (dvdx + dvdy + dvdz);   // a_iso_3d.c:680
}

// sources (a_iso_3d.c:689-693) then receivers for step it, one block
template <typename T>
__global__ void __launch_bounds__(256)
a3d_src_rcv(A3dDev<T> d, int nsrc, const int64_t *sidx, const int64_t *soff, const int64_t *slen,
            const T *sinc, int nrec, const int64_t *ridx, const int *rcomp, T *trace, int64_t it,
            int64_t row)
{
    if (threadIdx.x == 0) {
        for (int n = 0; n < nsrc; ++n)
            if (it < slen[n]) d.P[sidx[n]] += sinc[soff[n] + it];
    }
    __syncthreads();
    if (row < 0) return;
    for (int r = threadIdx.x; r < nrec; r += blockDim.x) {
        const int cp = rcomp[r];
        const T *f = cp == SM_VX ? d.vx : cp == SM_VY ? d.vy : cp == SM_VZ ? d.vz : d.P;
        trace[row * nrec + r] = f[ridx[r]];
    }
}

}  // namespace sm

using namespace sm;

struct sm_a3d {
    sm_ctx *ctx = nullptr;
    sm_a3d_params p{};
    int nx2 = 0, ny2 = 0, nz2 = 0;
    size_t nnn = 0;                 // cells per field (reference layout)
    int pz = 0;                     // device row pitch along z (cells)
    size_t nnp = 0;                 // cells per field on the device: nx2*ny2*pz
    size_t elt = 0;                 // bytes per stored state value
    size_t melt = 0;                // bytes per material / table value
    size_t telt = 0;                // bytes per recorded trace value
    bool fused = false;
    // device memory
    sm_dbuf fields[2];              // P, vx, vy, vz (ping-pong when fused)
    sm_dbuf mat;                    // rho, kdt
    sm_dbuf memv[2];                // mPx, mPy, mPz (ping-pong when fused)
    sm_dbuf memp;                   // mvx, mvy, mvz
    sm_dbuf tables;                 // 18 CPML tables
    sm_dbuf src;                    // cell, off, len (int64) | plane offsets (int) | increments
    sm_dbuf rcv;                    // idx (int64), comp (int)
    sm_dbuf trace;                  // [max_steps][nrec]
    sm_dbuf xpl;                    // per-plane x constants of the fused kernel
    std::vector<int> src_plane;     // sources of plane i: sorted entries [src_plane[i], [i+1])
    int cur = 0;                    // which ping-pong half holds the state
    size_t step = 0;
    size_t nsrc = 0, nrec = 0, max_steps = 0, rcv_step0 = 0;
    size_t src_plane_off = 0, src_inc_off = 0;   // byte offsets inside src
    int tj = 4, nchunks = 1;        // fused launch shape
    int variant = SM_A3D_CELL;      // fused kernel variant
    int sr = 0;                     // stochastic rounding of 16-bit stores
    int m16 = 0;                    // FP16 material (tile kernels, 16-bit state)
    float sR = 1.f, sK = 1.f;       // its scales: stored value * s = rho, K*dt
    int nblocks = 0;                // persistent grid (0: one block per item)
    int cus = 0;                    // compute units of the device
    // storage scales of the 16-bit formats (P, v, velocity-side and
    // pressure-side memory variables); 1 for FP32 and BF16
    float sP = 1.f, sV = 1.f, sMv = 1.f, sMp = 1.f;
    double rc_min = 0.0, d_min = 0.0;   // min rho*vp and min spacing, for the scales
    bool timing = false;
    double timed_ms = 0.0;
    size_t timed_launches = 0;
    std::vector<double> htab;       // FP64 CPML tables as built on the host
};

namespace {

// one set of x, y and z slab arrays, with rows of pz cells along z
size_t slab_cells(size_t nx2, size_t ny2, size_t pz, size_t np)
{
    return 2 * np * (ny2 * pz + nx2 * pz + nx2 * ny2);
}

bool uses_fused(const sm_a3d_params &p)
{
    if (p.storage == SM_STORE_F64) return false;
    if (p.storage == SM_STORE_F32) return !(p.flags & SM_A3D_NAIVE);
    return true;   // 16-bit storage exists only in the fused kernel
}

// device row pitch of the fused kernels: rows start on 128-byte cache
// lines (a multiple of 32 cells in FP32, 64 in 16-bit storage; every
// 64-cell wave access then covers whole lines) when that adds at most 5%
// (FP32) or 10% (16-bit) cells, otherwise a multiple of 8 cells (16 bytes
// of 16-bit storage; the TILE variants copy rows in 16-byte chunks).
// SM_A3D_PAD64 forces a multiple of 64 cells.  The two-kernel paths keep
// the reference layout.
size_t row_pitch(size_t nz2, const sm_a3d_params &p)
{
    if (!uses_fused(p)) return nz2;
    if (p.flags & SM_A3D_PAD64) return (nz2 + 63) / 64 * 64;
    const bool wide = p.storage == SM_STORE_F32;
    const size_t line = wide ? 32 : 64;
    const size_t pl = (nz2 + line - 1) / line * line;
    if ((pl - nz2) * (wide ? 20 : 10) <= nz2) return pl;
    return (nz2 + 7) / 8 * 8;
}

// reference-layout cell index -> device index
size_t dev_index(const sm_a3d *s, size_t idx)
{
    return idx / s->nz2 * s->pz + idx % s->nz2;
}

// rows of nz2 values from the host (reference layout) into rows of pz
// values on the device; the padding repeats each row's last value
template <typename T>
int upload_rows(void *dst, const double *src, size_t rows, size_t nz2, size_t pz, hipStream_t st)
{
    std::vector<T> tmp(rows * pz);
    for (size_t r = 0; r < rows; ++r) {
        for (size_t k = 0; k < nz2; ++k) tmp[r * pz + k] = (T)src[r * nz2 + k];
        for (size_t k = nz2; k < pz; ++k) tmp[r * pz + k] = (T)src[r * nz2 + nz2 - 1];
    }
    SM_HIP(hipMemcpyAsync(dst, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice, st));
    SM_HIP(hipStreamSynchronize(st));
    return SM_OK;
}

// Divides x by the power-of-two scale sc (returned) that puts its largest
// magnitude in [2^14, 2^15): FP16 then holds every value with 11 significant
// bits down to 2^-14 * sc, so a field whose nonzero magnitudes span more
// than 2^28 is refused rather than flushed towards zero.
int fp16_scale(std::vector<double> &x, const char *what, float &sc)
{
    double mx = 0.0, mn = 0.0;
    for (double v : x) {
        const double a = std::fabs(v);
        if (!std::isfinite(a)) return fail(SM_ERR_ARG, "16-bit material: %s is not finite", what);
        mx = std::max(mx, a);
        if (a > 0.0) mn = mn > 0.0 ? std::min(mn, a) : a;
    }
    if (mx == 0.0) { sc = 1.f; return SM_OK; }
    int e = 0;
    std::frexp(mx, &e);                       // mx = f * 2^e, f in [0.5, 1)
    const double scale = std::ldexp(1.0, e - 15);
    if (mn / scale < std::ldexp(1.0, -14))
        return fail(SM_ERR_ARG, "16-bit material: %s spans %g to %g, more than FP16 holds at "
                    "full precision under one scale", what, mn, mx);
    for (double &v : x) v /= scale;
    sc = (float)scale;
    return SM_OK;
}

// CPML tables of one axis (a_iso_3d.c:500-576), same write order.
void cpml_axis(size_t n2, size_t np, double d, double dt, double vmax, double order,
               double Rc, double kappa_max, double alpha_max, double *a, double *b, double *k,
               double *ah, double *bh, double *kh)
{
    for (size_t n = 0; n < n2; ++n) { a[n] = 0.0; b[n] = 1.0; k[n] = 1.0; }
    for (size_t n = 0; n + 1 < n2; ++n) { ah[n] = 0.0; bh[n] = 1.0; kh[n] = 1.0; }
    if (np == 0) return;
    const double da = alpha_max / np;
    const double L = np * d;
    const double d0 = -(order + 1.) * vmax * std::log(Rc) / (2. * L);
    for (size_t n = 0; n < np; ++n) {
        const double kn = 1. + (kappa_max - 1.) * std::pow(((np - n) * d / L), order);
        const double khn = 1. + (kappa_max - 1.) * std::pow(((np - n - 0.5) * d / L), order);
        k[n] = k[n2 - 1 - n] = kn;
        kh[n] = kh[n2 - 2 - n] = khn;
        const double alpha = n * da;
        const double alpha_h = (n + 0.5) * da;
        const double dn = d0 * std::pow(((np - n) * d / L), order);
        const double dnh = d0 * std::pow(((np - n - 0.5) * d / L), order);
        const double bn = std::exp(-(dn / k[n] + alpha) * dt);
        b[n] = b[n2 - 1 - n] = bn;
        const double bhn = std::exp(-(dnh / kh[n] + alpha_h) * dt);
        bh[n] = bh[n2 - 2 - n] = bhn;
        const double an = dn * (b[n] - 1.) / (k[n] * (dn + k[n] * alpha));
        a[n] = a[n2 - 1 - n] = an;
        const double ahn = dnh * (bh[n] - 1.) / (kh[n] * (dnh + kh[n] * alpha_h));
        ah[n] = ah[n2 - 2 - n] = ahn;
    }
}

template <typename T>
int upload(void *dst, const double *src, size_t n, hipStream_t st)
{
    std::vector<T> tmp(n);
    for (size_t i = 0; i < n; ++i) tmp[i] = (T)src[i];
    SM_HIP(hipMemcpyAsync(dst, tmp.data(), n * sizeof(T), hipMemcpyHostToDevice, st));
    SM_HIP(hipStreamSynchronize(st));
    return SM_OK;
}

// host conversions of the 16-bit storage formats
float bf16_to_float(uint16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

float half_to_float(uint16_t h)
{
    const uint32_t sign = (uint32_t)(h >> 15) << 31;
    uint32_t exp = (h >> 10) & 0x1f, man = h & 0x3ff, u;
    if (exp == 0) {
        if (man == 0) u = sign;
        else {  // subnormal: renormalise
            int e = -1;
            do { man <<= 1; ++e; } while (!(man & 0x400));
            u = sign | ((uint32_t)(127 - 15 - e) << 23) | ((man & 0x3ff) << 13);
        }
    } else if (exp == 31) u = sign | 0x7f800000u | (man << 13);
    else u = sign | ((exp + 127 - 15) << 23) | (man << 13);
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// state values on the host as double, scaled back to physical units
int state_to_host(sm_a3d *s, const void *dev, size_t n, size_t stride_elems, size_t count,
                  size_t pitch_elems, float scale, double *out)
{
    // copies `count` rows of n values, rows `pitch_elems` apart on the device
    const size_t e = s->elt;
    std::vector<unsigned char> tmp(n * count * e);
    SM_HIP(hipMemcpy2DAsync(tmp.data(), n * e, dev, pitch_elems * e, n * e, count,
                            hipMemcpyDeviceToHost, s->ctx->stream));
    SM_HIP(hipStreamSynchronize(s->ctx->stream));
    (void)stride_elems;
    const size_t tot = n * count;
    switch (s->p.storage) {
    case SM_STORE_F64:
        std::memcpy(out, tmp.data(), tot * 8);
        break;
    case SM_STORE_F32:
        for (size_t q = 0; q < tot; ++q) out[q] = ((const float *)tmp.data())[q];
        break;
    case SM_STORE_BF16:
        for (size_t q = 0; q < tot; ++q)
            out[q] = (double)bf16_to_float(((const uint16_t *)tmp.data())[q]) * scale;
        break;
    default:
        for (size_t q = 0; q < tot; ++q)
            out[q] = (double)half_to_float(((const uint16_t *)tmp.data())[q]) * scale;
        break;
    }
    return SM_OK;
}

template <typename T>
A3dDev<T> dev_view(sm_a3d *s, int half)
{
    A3dDev<T> d;
    d.nx2 = s->nx2; d.ny2 = s->ny2; d.nz2 = s->nz2; d.np = (int)s->p.np;
    T *f = (T *)s->fields[half].ptr;
    d.P = f; d.vx = f + s->nnp; d.vy = f + 2 * s->nnp; d.vz = f + 3 * s->nnp;
    const T *m = (const T *)s->mat.ptr;
    d.rho = m; d.kdt = m + s->nnp;
    const size_t np2 = 2 * s->p.np;
    T *mv = (T *)s->memv[half].ptr;
    d.mPx = mv; d.mPy = mv + np2 * s->ny2 * s->nz2; d.mPz = d.mPy + s->nx2 * np2 * s->nz2;
    T *mp = (T *)s->memp.ptr;
    d.mvx = mp; d.mvy = mp + np2 * s->ny2 * s->nz2; d.mvz = d.mvy + s->nx2 * np2 * s->nz2;
    return d;
}

template <typename T>
void axes(sm_a3d *s, CpmlAxis<T> ax[3])
{
    const T *t = (const T *)s->tables.ptr;
    const size_t n2[3] = {(size_t)s->nx2, (size_t)s->ny2, (size_t)s->nz2};
    for (int d = 0; d < 3; ++d) {
        ax[d].a = t; ax[d].b = t + n2[d]; ax[d].k = t + 2 * n2[d];
        ax[d].ah = t + 3 * n2[d]; ax[d].bh = t + 4 * n2[d]; ax[d].kh = t + 5 * n2[d];
        t += 6 * n2[d];
    }
}

// per-step salt of the stochastic rounding (splitmix64 finaliser of the step)
unsigned step_salt(size_t it)
{
    uint64_t z = (uint64_t)it + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (unsigned)(z ^ (z >> 31));
}

// kernel variant properties: planes prefetched (cell kernels); LDS tile
// kernels (z tiles of 64 cells starting at multiples of 64), with the CPML
// work split by tile, with non-temporal state stores
int variant_pf(int v) { return v == SM_A3D_CELL2 ? 2 : 1; }
bool variant_tile(int v) { return v >= SM_A3D_TILE && v <= SM_A3D_TILE_SPLIT_NT; }
bool variant_split(int v) { return v == SM_A3D_TILE_SPLIT || v == SM_A3D_TILE_SPLIT_NT; }
bool variant_nt(int v) { return v == SM_A3D_TILE_NT || v == SM_A3D_TILE_SPLIT_NT; }
bool variant_known(int v) { return v == SM_A3D_CELL || v == SM_A3D_CELL2 || variant_tile(v); }

// z tiles: 64 aligned cells (tile kernels), else 63 owned cells
int z_tiles(const sm_a3d *s)
{
    if (variant_tile(s->variant)) return (s->nz2 + 63) / 64;
    return (s->nz2 + 62) / 63;
}

// y tiles: rows_per_block rows per block
int y_tiles(const sm_a3d *s) { return (s->ny2 + s->tj - 1) / s->tj; }

int64_t rcv_row(const sm_a3d *s, size_t it)
{
    if (s->nrec == 0 || it < s->rcv_step0 || it - s->rcv_step0 >= s->max_steps) return -1;
    return (int64_t)(it - s->rcv_step0);
}

// per-launch timing with HIP events on the solver's stream
struct LaunchTimer {
    std::vector<hipEvent_t> ev;
    bool on = false;
    int begin(size_t n, bool enabled)
    {
        on = enabled;
        if (!on) return SM_OK;
        ev.resize(2 * n);
        for (auto &e : ev) SM_HIP(hipEventCreate(&e));
        return SM_OK;
    }
    int mark(hipStream_t st, size_t q)
    {
        if (on) SM_HIP(hipEventRecord(ev[q], st));
        return SM_OK;
    }
    int finish(sm_a3d *s)
    {
        if (!on) return SM_OK;
        SM_HIP(hipStreamSynchronize(s->ctx->stream));
        for (size_t q = 0; q + 1 < ev.size(); q += 2) {
            float ms = 0.f;
            SM_HIP(hipEventElapsedTime(&ms, ev[q], ev[q + 1]));
            s->timed_ms += ms;
            s->timed_launches += 1;
        }
        for (auto &e : ev) (void)hipEventDestroy(e);
        ev.clear();
        return SM_OK;
    }
};

template <typename T, bool EXACT>
int run_naive(sm_a3d *s, size_t nsteps)
{
    hipStream_t st = s->ctx->stream;
    A3dDev<T> d = dev_view<T>(s, s->cur);
    CpmlAxis<T> ax[3];
    axes<T>(s, ax);
    A3dConst<T> k;
    if constexpr (EXACT) {
        k.dx = s->p.dx; k.dy = s->p.dy; k.dz = s->p.dz; k.dt2 = s->p.dt;
    } else {
        k.dx = (T)(1.0 / s->p.dx); k.dy = (T)(1.0 / s->p.dy); k.dz = (T)(1.0 / s->p.dz);
        k.dt2 = (T)(2.0 * s->p.dt);
    }
    const dim3 blk(64, 4, 1);
    const dim3 grd((s->nz2 + 63) / 64, (s->ny2 + 3) / 4, s->nx2);
    const int64_t *sp = (const int64_t *)s->src.ptr;
    const T *sinc = (const T *)((const char *)s->src.ptr + s->src_inc_off);
    const int64_t *ridx = (const int64_t *)s->rcv.ptr;
    const int *rcomp = (const int *)(ridx + s->nrec);
    LaunchTimer tm;
    int rc = tm.begin(nsteps, s->timing);
    if (rc != SM_OK) return rc;
    for (size_t n = 0; n < nsteps; ++n) {
        const size_t it = s->step + n;
        if ((rc = tm.mark(st, 2 * n)) != SM_OK) return rc;
        a3d_vel_naive<T, EXACT><<<grd, blk, 0, st>>>(d, ax[0], ax[1], ax[2], k);
        a3d_pre_naive<T, EXACT><<<grd, blk, 0, st>>>(d, ax[0], ax[1], ax[2], k);
        if ((rc = tm.mark(st, 2 * n + 1)) != SM_OK) return rc;
        const int64_t row = rcv_row(s, it);
        if (s->nsrc || row >= 0)
            a3d_src_rcv<T><<<1, 256, 0, st>>>(d, (int)s->nsrc, sp, sp + s->nsrc, sp + 2 * s->nsrc,
                                              sinc, (int)s->nrec, ridx, rcomp, (T *)s->trace.ptr,
                                              (int64_t)it, row);
    }
    SM_HIP(hipGetLastError());
    s->step += nsteps;
    return tm.finish(s);
}

template <typename S>
FusedArgs<S> fused_args(sm_a3d *s)
{
    FusedArgs<S> a{};
    a.nx2 = s->nx2; a.ny2 = s->ny2; a.nz2 = s->nz2; a.np = (int)s->p.np;
    a.ntk = z_tiles(s);
    a.ntiles = a.ntk * y_tiles(s);
    a.chunk_len = (s->nx2 + s->nchunks - 1) / s->nchunks;
    a.nreal = a.ntiles * s->nchunks;
    a.nitems = (a.nreal + 7) / 8 * 8;
    a.nblocks = s->nblocks > 0 ? std::min(s->nblocks, a.nitems) : a.nitems;
    a.nnn = (unsigned)s->nnp;
    a.pz = s->pz;
    const unsigned np2 = 2u * (unsigned)s->p.np;
    a.memp = (S *)s->memp.ptr;
    a.my = np2 * (unsigned)s->ny2 * (unsigned)s->pz;
    a.mz = a.my + (unsigned)s->nx2 * np2 * (unsigned)s->pz;
    a.mat = (const float *)s->mat.ptr;
    a.sR = s->sR; a.sK = s->sK;
    a.xpl = (const XPlane *)s->xpl.ptr;
    a.tab = (const float *)s->tables.ptr;
    a.toff_y = 6 * s->nx2;
    a.toff_z = 6 * (s->nx2 + s->ny2);
    a.idx = (float)(1.0 / s->p.dx); a.idy = (float)(1.0 / s->p.dy); a.idz = (float)(1.0 / s->p.dz);
    a.two_dt = (float)(2.0 * s->p.dt);
    a.sP = s->sP; a.isP = 1.f / s->sP; a.sV = s->sV; a.isV = 1.f / s->sV;
    a.sMv = s->sMv; a.isMv = 1.f / s->sMv; a.sMp = s->sMp; a.isMp = 1.f / s->sMp;
    const char *sb = (const char *)s->src.ptr;
    a.src = (const int64_t *)sb;
    a.src_inc = (const float *)(sb + s->src_inc_off);
    a.nsrc = (int)s->nsrc;
    a.nrec = (int)s->nrec;
    a.ridx = (const int64_t *)s->rcv.ptr;
    a.rcomp = (const int *)(a.ridx + s->nrec);
    a.trace = (float *)s->trace.ptr;
    return a;
}

template <typename S>
void fused_bind(sm_a3d *s, FusedArgs<S> &a, int in)
{
    a.in = (const S *)s->fields[in].ptr;
    a.out = (S *)s->fields[1 - in].ptr;
    a.mem_in = (const S *)s->memv[in].ptr;
    a.mem_out = (S *)s->memv[1 - in].ptr;
}

template <typename S>
int run_fused(sm_a3d *s, size_t nsteps)
{
    hipStream_t st = s->ctx->stream;
    FusedArgs<S> a = fused_args<S>(s);
    const int per_gather = 64 * s->tj;
    LaunchTimer tm;
    int rc = tm.begin(nsteps, s->timing);
    if (rc != SM_OK) return rc;
    for (size_t n = 0; n < nsteps; ++n) {
        const size_t it = s->step + n;
        fused_bind<S>(s, a, s->cur);
        a.step = (int64_t)it;
        a.salt = step_salt(it);
        a.sr = s->sr;
        // receivers of the previous step, unless an earlier run() already did
        a.gather_row = (n > 0 && it > 0) ? rcv_row(s, it - 1) : -1;
        a.gather_blocks = a.gather_row >= 0 ? (int)((s->nrec + per_gather - 1) / per_gather) : 0;
        if ((rc = tm.mark(st, 2 * n)) != SM_OK) return rc;
        if ((rc = variant_tile(s->variant)
                      ? tile_launch<S>(a, s->tj, variant_split(s->variant), variant_nt(s->variant),
                                       s->m16 != 0, st)
                      : fused_launch<S>(a, s->tj, variant_pf(s->variant), st)) != SM_OK)
            return fail(rc, "fused launch failed (tj=%d): %s", s->tj,
                        hipGetErrorString(hipGetLastError()));
        if ((rc = tm.mark(st, 2 * n + 1)) != SM_OK) return rc;
        s->cur = 1 - s->cur;
    }
    // receivers of the last step, from the state it produced
    fused_bind<S>(s, a, s->cur);
    a.gather_row = rcv_row(s, s->step + nsteps - 1);
    if ((rc = gather_launch<S>(a, st)) != SM_OK) return fail(rc, "gather launch failed");
    s->step += nsteps;
    return tm.finish(s);
}

int zero_state(sm_a3d *s)
{
    hipStream_t st = s->ctx->stream;
    for (int h = 0; h < 2; ++h) {
        if (s->fields[h].ptr) SM_HIP(hipMemsetAsync(s->fields[h].ptr, 0, s->fields[h].bytes, st));
        if (s->memv[h].ptr) SM_HIP(hipMemsetAsync(s->memv[h].ptr, 0, s->memv[h].bytes, st));
    }
    if (s->memp.ptr) SM_HIP(hipMemsetAsync(s->memp.ptr, 0, s->memp.bytes, st));
    if (s->trace.ptr) SM_HIP(hipMemsetAsync(s->trace.ptr, 0, s->trace.bytes, st));
    SM_HIP(hipStreamSynchronize(st));
    s->cur = 0;
    s->step = 0;
    s->rcv_step0 = 0;
    return SM_OK;
}

void release(sm_a3d *s)
{
    for (int h = 0; h < 2; ++h) { dfree(s->fields[h]); dfree(s->memv[h]); }
    dfree(s->mat); dfree(s->memp); dfree(s->tables); dfree(s->src); dfree(s->rcv); dfree(s->trace);
    dfree(s->xpl);
}

// per-plane x records of the fused kernel from the FP64 tables and the
// per-plane source ranges
int upload_xplanes(sm_a3d *s)
{
    if (!s->fused) return SM_OK;
    const int nx2 = s->nx2;
    const double *t = s->htab.data();   // x axis: a, b, k, ah, bh, kh (nx2 each)
    std::vector<XPlane> h(nx2);
    for (int i = 0; i < nx2; ++i) {
        XPlane &x = h[i];
        std::memset(&x, 0, sizeof(x));
        if (i < nx2 - 1) {
            x.ah = (float)t[3 * nx2 + i];
            x.bh = (float)t[4 * nx2 + i];
            x.ikh = (float)(1.0 / t[5 * nx2 + i]);
        } else {
            x.ah = 0.f; x.bh = 1.f; x.ikh = 1.f;
        }
        x.a = (float)t[i];
        x.b = (float)t[nx2 + i];
        x.ik = (float)(1.0 / t[2 * nx2 + i]);
        x.sxv = slab_vel(i, nx2, (int)s->p.np);
        x.sxp = slab_pre(i, nx2, (int)s->p.np);
        x.q0 = s->src_plane.empty() ? 0 : s->src_plane[i];
        x.q1 = s->src_plane.empty() ? 0 : s->src_plane[i + 1];
    }
    if (!s->xpl.ptr) {
        int rc = dalloc(s->xpl, nx2 * sizeof(XPlane));
        if (rc != SM_OK) return rc;
    }
    SM_HIP(hipMemcpyAsync(s->xpl.ptr, h.data(), nx2 * sizeof(XPlane), hipMemcpyHostToDevice,
                          s->ctx->stream));
    SM_HIP(hipStreamSynchronize(s->ctx->stream));
    return SM_OK;
}

size_t elt_size(int storage)
{
    switch (storage) {
    case SM_STORE_F32: return 4;
    case SM_STORE_F64: return 8;
    case SM_STORE_BF16: return 2;
    case SM_STORE_F16: return 2;
    default: return 0;
    }
}


int resident_blocks(sm_a3d *s, int tj)
{
    int per_cu = 0, rc = SM_OK;
    const bool tile = variant_tile(s->variant);
    const bool split = variant_split(s->variant), nt = variant_nt(s->variant), sr = s->sr != 0;
    const bool m16 = s->m16 != 0;
    const int pf = variant_pf(s->variant);
    switch (s->p.storage) {
    case SM_STORE_F32:
        rc = tile ? tile_max_blocks_per_cu<float>(tj, split, nt, sr, m16, &per_cu)
                  : fused_max_blocks_per_cu<float>(tj, pf, &per_cu);
        break;
    case SM_STORE_BF16:
        rc = tile ? tile_max_blocks_per_cu<__hip_bfloat16>(tj, split, nt, sr, m16, &per_cu)
                  : fused_max_blocks_per_cu<__hip_bfloat16>(tj, pf, &per_cu);
        break;
    default:
        rc = tile ? tile_max_blocks_per_cu<__half>(tj, split, nt, sr, m16, &per_cu)
                  : fused_max_blocks_per_cu<__half>(tj, pf, &per_cu);
        break;
    }
    if (rc != SM_OK || per_cu < 1) per_cu = 1;
    return per_cu * std::max(s->cus, 1);
}

// chunk count with no empty chunk
int valid_chunks(int nx2, int want)
{
    want = std::max(1, std::min(want, nx2));
    const int len = (nx2 + want - 1) / want;
    return (nx2 + len - 1) / len;
}

// default launch shape: 4 waves per block; x chunks so that every resident
// block slot gets about three items; persistent grid of resident size
void default_launch(sm_a3d *s)
{
    s->tj = 4;
    const int ntiles = z_tiles(s) * y_tiles(s);
    const int slots = resident_blocks(s, s->tj);
    int nch = (3 * slots + ntiles - 1) / ntiles;
    nch = std::min(nch, std::max(1, s->nx2 / 24));
    s->nchunks = valid_chunks(s->nx2, nch);
    s->nblocks = slots;
}

}  // namespace

extern "C" {

size_t sm_a3d_bytes_needed(const sm_a3d_params *p, size_t nrec, size_t max_steps)
{
    if (!p) return 0;
    const size_t e = elt_size(p->storage);
    if (!e) return 0;
    const size_t me = p->storage == SM_STORE_F64 ? 8 : 4;
    const size_t mat_e = (p->flags & SM_A3D_MAT16) ? 2 : me;
    const size_t nx2 = p->nx + 2 * p->np, ny2 = p->ny + 2 * p->np, nz2 = p->nz + 2 * p->np;
    const size_t nnn = nx2 * ny2 * row_pitch(nz2, *p);
    const size_t halves = uses_fused(*p) ? 2 : 1;
    const size_t slabs = slab_cells(nx2, ny2, row_pitch(nz2, *p), p->np);
    return e * (halves * 4 * nnn + (halves + 1) * slabs) + mat_e * 2 * nnn +
           me * 6 * (nx2 + ny2 + nz2) + me * nrec * max_steps + 16 * nrec + 4 * nx2;
}

int sm_a3d_create(sm_ctx *ctx, const sm_a3d_params *pp, const double *vp, const double *rho,
                  sm_a3d **out)
{
    SM_CHECK(ctx && pp && vp && rho && out, SM_ERR_ARG, "sm_a3d_create: NULL argument");
    const sm_a3d_params &p = *pp;
    SM_CHECK(p.nx >= 1 && p.ny >= 1 && p.nz >= 1, SM_ERR_ARG, "sm_a3d_create: empty grid");
    SM_CHECK(p.dx > 0 && p.dy > 0 && p.dz > 0 && p.dt > 0, SM_ERR_ARG,
             "sm_a3d_create: spacings and dt must be > 0");
    SM_CHECK(p.dx == p.dz, SM_ERR_ARG,
             "grid cell size must be equal in X and Z (a_iso_3d.c:298-301)");
    const size_t e = elt_size(p.storage);
    SM_CHECK(e != 0, SM_ERR_ARG, "sm_a3d_create: storage %d not supported", p.storage);
    SM_CHECK((p.flags & ~(SM_A3D_NAIVE | SM_A3D_SROUND | SM_A3D_PAD64 | SM_A3D_MAT16)) == 0,
             SM_ERR_ARG, "sm_a3d_create: unknown flags %#x", p.flags);
    SM_CHECK(!(p.flags & SM_A3D_MAT16) ||
                 ((p.storage == SM_STORE_BF16 || p.storage == SM_STORE_F16) && !(p.flags & SM_A3D_SROUND)),
             SM_ERR_ARG, "sm_a3d_create: SM_A3D_MAT16 needs BF16 or FP16 storage with "
             "round-to-nearest stores (no SM_A3D_SROUND)");
    const size_t nx2 = p.nx + 2 * p.np, ny2 = p.ny + 2 * p.np, nz2 = p.nz + 2 * p.np;
    SM_CHECK(nx2 < 65536 && ny2 < 65536 && nz2 < 65536, SM_ERR_ARG, "sm_a3d_create: grid too large");
    const size_t nnn = nx2 * ny2 * nz2;
    SM_CHECK(nnn < ((size_t)1 << 34), SM_ERR_ARG, "sm_a3d_create: grid too large");

    // a_iso_3d.c:368-374: Vp maximum over the first nx*nz padded entries
    double vpmax = 0.0;
    for (size_t i = 0; i < p.nx; ++i)
        for (size_t j = 0; j < p.nz; ++j) {
            const double v = vp[i * p.nz + j];
            vpmax = v > vpmax ? v : vpmax;
        }
    // a_iso_3d.c:411-417
    const double lim = p.dt * vpmax *
                       std::sqrt(1. / (p.dx * p.dx) + 1. / (p.dy * p.dy) + 1. / (p.dz * p.dz));
    if (lim > 1.)
        return fail(SM_ERR_CFL, "time step too large: dt*vpmax*sqrt(sum 1/d^2) = %f > 1", lim);
    for (size_t n = 0; n < nnn; ++n)
        SM_CHECK(rho[n] > 0.0 && std::isfinite(vp[n]), SM_ERR_ARG,
                 "sm_a3d_create: rho must be > 0 and vp finite (cell %zu)", n);

    SM_CHECK(set_device(ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    sm_a3d *s = new sm_a3d;
    s->ctx = ctx;
    s->p = p;
    s->nx2 = (int)nx2; s->ny2 = (int)ny2; s->nz2 = (int)nz2;
    s->nnn = nnn;
    s->elt = e;
    s->melt = p.storage == SM_STORE_F64 ? 8 : 4;
    s->telt = s->melt;
    s->fused = uses_fused(p);
    s->pz = (int)row_pitch(nz2, p);
    s->nnp = nx2 * ny2 * (size_t)s->pz;
    const size_t nnp = s->nnp, pz = (size_t)s->pz;
    s->sr = (p.flags & SM_A3D_SROUND) && (p.storage == SM_STORE_BF16 || p.storage == SM_STORE_F16);
    s->m16 = (p.flags & SM_A3D_MAT16) != 0;
    if (s->m16) s->variant = SM_A3D_TILE_SPLIT;   // the cell kernels read FP32 material
    // the fused kernel addresses the four state fields of one buffer and
    // the two material fields with 32-bit byte offsets
    if (s->fused && (4 * nnp * e > (size_t)0xFFFFFF00u || 2 * nnp * 4 > (size_t)0xFFFFFF00u)) {
        delete s;
        return fail(SM_ERR_ARG, "sm_a3d_create: %zu cells (rows padded to %zu) exceed the fused "
                    "kernel's 32-bit offsets (limit 2^28 cells with FP32 storage, 2^29 with "
                    "16-bit); use SM_A3D_NAIVE with FP32 storage", nnn, pz);
    }
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess)
            s->cus = prop.multiProcessorCount;
    }
    if (s->fused) default_launch(s);
    // one set (velocity or pressure) of x, y and z slab arrays
    const size_t slab_set = slab_cells(nx2, ny2, pz, p.np);
    int rc = SM_OK;
    const int halves = s->fused ? 2 : 1;
    for (int h = 0; h < halves && rc == SM_OK; ++h) {
        rc = dalloc(s->fields[h], 4 * nnp * e);
        if (rc == SM_OK) rc = dalloc(s->memv[h], slab_set * e);
    }
    if (rc == SM_OK) rc = dalloc(s->mat, 2 * nnp * (s->m16 ? 2 : s->melt));
    if (rc == SM_OK) rc = dalloc(s->memp, slab_set * e);
    if (rc == SM_OK) rc = dalloc(s->tables, 6 * (nx2 + ny2 + nz2) * s->melt);
    if (rc != SM_OK) { release(s); delete s; return rc; }
    {
        // extent of the slab set as dev_view lays it out: must end inside
        // the allocation (x: 2np*ny2*pz, y: nx2*2np*pz, z: nx2*ny2*2np)
        const size_t np2 = 2 * p.np;
        const size_t end = np2 * ny2 * pz + nx2 * np2 * pz + nx2 * ny2 * np2;
        if (end * e > s->memp.bytes || end * e > s->memv[0].bytes) {
            release(s); delete s;
            return fail(SM_ERR_STATE, "internal: slab layout exceeds its allocation");
        }
    }

    // material: rho as read, K*dt = rho vp^2 dt (a_iso_3d.c:407-409)
    {
        std::vector<double> kdt(nnn);
        for (size_t n = 0; n < nnn; ++n) kdt[n] = rho[n] * vp[n] * vp[n] * p.dt;
        hipStream_t st = ctx->stream;
        const size_t rows = nx2 * ny2;
        if (p.storage == SM_STORE_F64) {
            rc = upload_rows<double>(s->mat.ptr, rho, rows, nz2, pz, st);
            if (rc == SM_OK)
                rc = upload_rows<double>((double *)s->mat.ptr + nnp, kdt.data(), rows, nz2, pz, st);
        } else if (s->m16) {
            // FP16 with one power-of-two scale per field, rounded once from FP64
            std::vector<double> r16(rho, rho + nnn);
            rc = fp16_scale(r16, "rho", s->sR);
            if (rc == SM_OK) rc = fp16_scale(kdt, "K*dt", s->sK);
            if (rc == SM_OK) rc = upload_rows<_Float16>(s->mat.ptr, r16.data(), rows, nz2, pz, st);
            if (rc == SM_OK)
                rc = upload_rows<_Float16>((_Float16 *)s->mat.ptr + nnp, kdt.data(), rows, nz2, pz, st);
        } else {
            rc = upload_rows<float>(s->mat.ptr, rho, rows, nz2, pz, st);
            if (rc == SM_OK)
                rc = upload_rows<float>((float *)s->mat.ptr + nnp, kdt.data(), rows, nz2, pz, st);
        }
        double rcm = 1e300;
        for (size_t n = 0; n < nnn; ++n) rcm = std::min(rcm, rho[n] * std::max(vp[n], 1.0));
        s->rc_min = rcm;
        s->d_min = std::min(p.dx, std::min(p.dy, p.dz));
    }
    // CPML tables (a_iso_3d.c:525-576)
    if (rc == SM_OK) {
        const double pi = 4.0 * std::atan(1.0);
        double alpha_max = pi * p.f0;
        if (p.alpha_max == 0) alpha_max = 0.0;
        const size_t n2[3] = {nx2, ny2, nz2};
        const double dd[3] = {p.dx, p.dy, p.dz};
        s->htab.assign(6 * (nx2 + ny2 + nz2), 0.0);
        double *t = s->htab.data();
        for (int d = 0; d < 3; ++d) {
            cpml_axis(n2[d], p.np, dd[d], p.dt, p.vmax, p.pml_order, p.Rc, p.kappa_max, alpha_max,
                      t, t + n2[d], t + 2 * n2[d], t + 3 * n2[d], t + 4 * n2[d], t + 5 * n2[d]);
            t += 6 * n2[d];
        }
        if (p.storage == SM_STORE_F64) {
            rc = upload<double>(s->tables.ptr, s->htab.data(), s->htab.size(), ctx->stream);
        } else {
            // FP32 paths store 1/kappa in place of kappa
            std::vector<double> inv(s->htab);
            double *q = inv.data();
            for (int d = 0; d < 3; ++d) {
                for (size_t n = 0; n < n2[d]; ++n) q[2 * n2[d] + n] = 1.0 / q[2 * n2[d] + n];
                for (size_t n = 0; n + 1 < n2[d]; ++n) q[5 * n2[d] + n] = 1.0 / q[5 * n2[d] + n];
                q += 6 * n2[d];
            }
            rc = upload<float>(s->tables.ptr, inv.data(), inv.size(), ctx->stream);
        }
    }
    if (rc == SM_OK) rc = zero_state(s);
    // an empty source set still needs the per-plane offsets of the fused path
    if (rc == SM_OK) rc = sm_a3d_set_sources(s, 0, nullptr, nullptr, nullptr);
    if (rc != SM_OK) { release(s); delete s; return rc; }
    *out = s;
    return SM_OK;
}

int sm_a3d_destroy(sm_a3d *s)
{
    if (!s) return SM_OK;
    (void)set_device(s->ctx);
    (void)hipStreamSynchronize(s->ctx->stream);
    release(s);
    delete s;
    return SM_OK;
}

int sm_a3d_set_sources(sm_a3d *s, size_t nsrc, const size_t *idx, const size_t *length,
                       const double *fct)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_a3d_set_sources: NULL solver");
    SM_CHECK(nsrc == 0 || (idx && length && fct), SM_ERR_ARG, "sm_a3d_set_sources: NULL array");
    SM_CHECK(nsrc < (1u << 20), SM_ERR_ARG, "sm_a3d_set_sources: too many sources");
    size_t total = 0;
    for (size_t n = 0; n < nsrc; ++n) {
        SM_CHECK(idx[n] < s->nnn, SM_ERR_ARG, "source %zu: index %zu outside the grid", n, idx[n]);
        total += length[n];
    }
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    SM_HIP(hipStreamSynchronize(s->ctx->stream));
    dfree(s->src);
    s->nsrc = 0;
    // sources in cell order, stable, so several sources on one cell add in
    // the reference's order (a_iso_3d.c:689-693)
    std::vector<size_t> order(nsrc), offs(nsrc);
    size_t off = 0;
    for (size_t n = 0; n < nsrc; ++n) { order[n] = n; offs[n] = off; off += length[n]; }
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return idx[a] < idx[b]; });
    const size_t plane = (size_t)s->ny2 * s->pz;
    const size_t head = 3 * nsrc * sizeof(int64_t);
    const size_t plane_bytes = ((size_t)s->nx2 + 1) * sizeof(int);
    s->src_plane_off = head;
    s->src_inc_off = (head + plane_bytes + 7) / 8 * 8;
    const size_t bytes = s->src_inc_off + total * s->melt;
    int rc = dalloc(s->src, bytes);
    if (rc != SM_OK) return rc;
    std::vector<unsigned char> h(bytes, 0);
    int64_t *hi = (int64_t *)h.data();
    int *hp = (int *)(h.data() + s->src_plane_off);
    size_t q = 0;
    for (size_t r = 0; r < nsrc; ++r) {
        const size_t n = order[r];
        hi[r] = (int64_t)dev_index(s, idx[n]);
        hi[nsrc + r] = (int64_t)q;
        hi[2 * nsrc + r] = (int64_t)length[n];
        // increment dt*fct*1e6 (a_iso_3d.c:691), rounded once to the working type
        for (size_t t = 0; t < length[n]; ++t, ++q) {
            const double inc = s->p.dt * fct[offs[n] + t] * 1.e6;
            if (s->melt == 8) ((double *)(h.data() + s->src_inc_off))[q] = inc;
            else ((float *)(h.data() + s->src_inc_off))[q] = (float)inc;
        }
    }
    s->src_plane.assign(s->nx2 + 1, 0);
    for (int i = 0, r = 0; i <= s->nx2; ++i) {
        while (r < (int)nsrc && (size_t)hi[r] / plane < (size_t)i) ++r;
        hp[i] = r;
        s->src_plane[i] = r;
    }
    SM_HIP(hipMemcpyAsync(s->src.ptr, h.data(), bytes, hipMemcpyHostToDevice, s->ctx->stream));
    SM_HIP(hipStreamSynchronize(s->ctx->stream));
    s->nsrc = nsrc;
    // FP16 storage: power-of-two scales from the injected pressure, while the
    // state is still zero (a rescale of a live state would be needed later).
    // The largest pressure is bounded by the total injected at one cell; the
    // velocities, dP/dx and dv/dx scale with it.  Targets sit 8x below the
    // FP16 maximum (65504).
    if (s->p.storage == SM_STORE_F16 && s->step == 0 && nsrc > 0) {
        double pmax = 0.0;
        for (size_t n = 0; n < nsrc; ++n) {
            double acc = 0.0;
            for (size_t t = 0; t < length[n]; ++t) acc += std::fabs(s->p.dt * fct[offs[n] + t] * 1.e6);
            pmax = std::max(pmax, acc);
        }
        if (pmax > 0.0) {
            auto p2 = [](double v) { return std::ldexp(1.0f, (int)std::ceil(std::log2(v / 8192.0))); };
            const double vmax = pmax / s->rc_min;
            s->sP = p2(pmax);
            s->sV = p2(vmax);
            s->sMv = p2(2.0 * pmax / s->d_min);
            s->sMp = p2(2.0 * vmax / s->d_min);
        }
    }
    return upload_xplanes(s);
}

int sm_a3d_set_receivers(sm_a3d *s, size_t nrec, const size_t *idx, const int *comp,
                         size_t max_steps)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_a3d_set_receivers: NULL solver");
    SM_CHECK(nrec == 0 || (idx && comp), SM_ERR_ARG, "sm_a3d_set_receivers: NULL array");
    SM_CHECK(nrec < (1u << 24), SM_ERR_ARG, "sm_a3d_set_receivers: too many receivers");
    for (size_t r = 0; r < nrec; ++r) {
        SM_CHECK(idx[r] < s->nnn, SM_ERR_ARG, "receiver %zu: index %zu outside the grid", r, idx[r]);
        SM_CHECK(comp[r] == SM_VX || comp[r] == SM_VY || comp[r] == SM_VZ || comp[r] == SM_P,
                 SM_ERR_ARG, "receiver %zu: component %d not recordable in 3D", r, comp[r]);
    }
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    SM_HIP(hipStreamSynchronize(s->ctx->stream));
    dfree(s->rcv);
    dfree(s->trace);
    s->nrec = 0;
    s->max_steps = 0;
    if (nrec == 0 || max_steps == 0) return SM_OK;
    int rc = dalloc(s->rcv, nrec * (sizeof(int64_t) + sizeof(int)));
    if (rc == SM_OK) rc = dalloc(s->trace, nrec * max_steps * s->telt);
    if (rc != SM_OK) { dfree(s->rcv); dfree(s->trace); return rc; }
    std::vector<unsigned char> h(nrec * (sizeof(int64_t) + sizeof(int)));
    int64_t *hi = (int64_t *)h.data();
    int *hc = (int *)(hi + nrec);
    for (size_t r = 0; r < nrec; ++r) { hi[r] = (int64_t)dev_index(s, idx[r]); hc[r] = comp[r]; }
    SM_HIP(hipMemcpyAsync(s->rcv.ptr, h.data(), h.size(), hipMemcpyHostToDevice, s->ctx->stream));
    SM_HIP(hipMemsetAsync(s->trace.ptr, 0, s->trace.bytes, s->ctx->stream));
    SM_HIP(hipStreamSynchronize(s->ctx->stream));
    s->nrec = nrec;
    s->max_steps = max_steps;
    s->rcv_step0 = s->step;
    return SM_OK;
}

int sm_a3d_set_launch(sm_a3d *s, int rows_per_block, int x_chunks)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_a3d_set_launch: NULL solver");
    SM_CHECK(s->fused, SM_ERR_STATE, "sm_a3d_set_launch: only the fused kernel has a launch shape");
    SM_CHECK(rows_per_block == 1 || rows_per_block == 2 || rows_per_block == 4 ||
                 rows_per_block == 8, SM_ERR_ARG, "rows_per_block must be 1, 2, 4 or 8");
    SM_CHECK(!variant_nt(s->variant) || rows_per_block >= 4, SM_ERR_ARG,
             "the non-temporal tile variants take rows_per_block 4 or 8");
    SM_CHECK(!s->m16 || rows_per_block >= 4, SM_ERR_ARG,
             "16-bit material (SM_A3D_MAT16) takes rows_per_block 4 or 8");
    SM_CHECK(x_chunks >= 1 && x_chunks <= s->nx2, SM_ERR_ARG, "x_chunks must be in [1, nx2]");
    // every chunk must own at least one plane
    const int len = (s->nx2 + x_chunks - 1) / x_chunks;
    SM_CHECK((x_chunks - 1) * len < s->nx2, SM_ERR_ARG, "x_chunks %d leaves an empty chunk", x_chunks);
    s->tj = rows_per_block;
    s->nchunks = x_chunks;
    s->nblocks = resident_blocks(s, rows_per_block);
    return SM_OK;
}

int sm_a3d_set_variant(sm_a3d *s, int variant)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_a3d_set_variant: NULL solver");
    SM_CHECK(s->fused, SM_ERR_STATE, "sm_a3d_set_variant: only the fused kernel has variants");
    SM_CHECK(variant_known(variant), SM_ERR_ARG, "sm_a3d_set_variant: unknown variant %d", variant);
    SM_CHECK(!s->m16 || variant_tile(variant), SM_ERR_ARG,
             "sm_a3d_set_variant: 16-bit material (SM_A3D_MAT16) runs on the tile variants only");
    SM_CHECK(!variant_tile(variant) || s->pz % 8 == 0, SM_ERR_ARG,
             "sm_a3d_set_variant: the tile kernel needs a row pitch of a multiple of 8 cells");
    if (variant == s->variant) return SM_OK;
    s->variant = variant;
    default_launch(s);
    return SM_OK;
}

int sm_a3d_get_variant(const sm_a3d *s, int *variant)
{
    SM_CHECK(s && variant, SM_ERR_ARG, "sm_a3d_get_variant: NULL argument");
    *variant = s->fused ? s->variant : -1;
    return SM_OK;
}

int sm_a3d_get_launch(const sm_a3d *s, int *rows_per_block, int *x_chunks)
{
    SM_CHECK(s && rows_per_block && x_chunks, SM_ERR_ARG, "sm_a3d_get_launch: NULL argument");
    *rows_per_block = s->fused ? s->tj : 0;
    *x_chunks = s->fused ? s->nchunks : 0;
    return SM_OK;
}

int sm_a3d_set_timing(sm_a3d *s, int on)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_a3d_set_timing: NULL solver");
    s->timing = on != 0;
    s->timed_ms = 0.0;
    s->timed_launches = 0;
    return SM_OK;
}

int sm_a3d_get_timing(const sm_a3d *s, double *ms_total, size_t *launches)
{
    SM_CHECK(s && ms_total && launches, SM_ERR_ARG, "sm_a3d_get_timing: NULL argument");
    *ms_total = s->timed_ms;
    *launches = s->timed_launches;
    return SM_OK;
}

int sm_a3d_run(sm_a3d *s, size_t nsteps)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_a3d_run: NULL solver");
    if (nsteps == 0) return SM_OK;
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    switch (s->p.storage) {
    case SM_STORE_F64: return run_naive<double, true>(s, nsteps);
    case SM_STORE_F32:
        return s->fused ? run_fused<float>(s, nsteps) : run_naive<float, false>(s, nsteps);
    case SM_STORE_BF16: return run_fused<__hip_bfloat16>(s, nsteps);
    case SM_STORE_F16: return run_fused<__half>(s, nsteps);
    default: return fail(SM_ERR_STATE, "sm_a3d_run: bad storage");
    }
}

int sm_a3d_current_step(const sm_a3d *s, size_t *it)
{
    SM_CHECK(s && it, SM_ERR_ARG, "sm_a3d_current_step: NULL argument");
    *it = s->step;
    return SM_OK;
}

int sm_a3d_get_traces(sm_a3d *s, size_t it0, size_t count, double *out)
{
    SM_CHECK(s && (out || count == 0), SM_ERR_ARG, "sm_a3d_get_traces: NULL argument");
    SM_CHECK(it0 + count <= s->max_steps, SM_ERR_ARG,
             "sm_a3d_get_traces: steps [%zu,%zu) beyond the %zu recorded", it0, it0 + count,
             s->max_steps);
    if (count == 0 || s->nrec == 0) return SM_OK;
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    const size_t n = count * s->nrec;
    std::vector<unsigned char> tmp(n * s->telt);
    SM_HIP(hipMemcpyAsync(tmp.data(), (const char *)s->trace.ptr + it0 * s->nrec * s->telt,
                          n * s->telt, hipMemcpyDeviceToHost, s->ctx->stream));
    SM_HIP(hipStreamSynchronize(s->ctx->stream));
    if (s->telt == 8) std::memcpy(out, tmp.data(), n * 8);
    else for (size_t q = 0; q < n; ++q) out[q] = ((const float *)tmp.data())[q];
    return SM_OK;
}

static int field_slot(int comp)
{
    return comp == SM_P ? 0 : comp == SM_VX ? 1 : comp == SM_VY ? 2 : comp == SM_VZ ? 3 : -1;
}

int sm_a3d_get_field(sm_a3d *s, int comp, double *out)
{
    SM_CHECK(s && out, SM_ERR_ARG, "sm_a3d_get_field: NULL argument");
    const int slot = field_slot(comp);
    SM_CHECK(slot >= 0, SM_ERR_ARG, "sm_a3d_get_field: component %d", comp);
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    const char *src = (const char *)s->fields[s->cur].ptr + slot * s->nnp * s->elt;
    return state_to_host(s, src, s->nz2, 0, (size_t)s->nx2 * s->ny2, s->pz,
                         slot == 0 ? s->sP : s->sV, out);
}

int sm_a3d_get_slice_xz(sm_a3d *s, int comp, size_t iy, double *out)
{
    SM_CHECK(s && out, SM_ERR_ARG, "sm_a3d_get_slice_xz: NULL argument");
    SM_CHECK(iy < (size_t)s->ny2, SM_ERR_ARG, "sm_a3d_get_slice_xz: iy %zu >= ny2", iy);
    const int slot = field_slot(comp);
    SM_CHECK(slot >= 0, SM_ERR_ARG, "sm_a3d_get_slice_xz: component %d", comp);
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    const char *base = (const char *)s->fields[s->cur].ptr + slot * s->nnp * s->elt +
                       iy * (size_t)s->pz * s->elt;
    return state_to_host(s, base, s->nz2, 0, s->nx2, (size_t)s->ny2 * s->pz,
                         slot == 0 ? s->sP : s->sV, out);
}

int sm_a3d_reset(sm_a3d *s)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_a3d_reset: NULL solver");
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    return zero_state(s);
}

int sm_a3d_device_bytes(const sm_a3d *s, size_t *bytes)
{
    SM_CHECK(s && bytes, SM_ERR_ARG, "sm_a3d_device_bytes: NULL argument");
    size_t b = s->mat.bytes + s->memp.bytes + s->tables.bytes + s->src.bytes + s->rcv.bytes +
               s->trace.bytes;
    for (int h = 0; h < 2; ++h) b += s->fields[h].bytes + s->memv[h].bytes;
    *bytes = b;
    return SM_OK;
}

}  // extern "C"
