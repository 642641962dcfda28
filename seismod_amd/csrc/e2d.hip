// e2d.hip -- 2D elastic split-PML solver (replaces e_iso.c:423-500, 565-670,
// propagate.c:2511-2670, pml.c:39-1403, src.c:170-184).
//
// Device layout: one padded grid of NX = nx+2Np by NZ = nz+2Np cells,
// index I*NZ+J (z fastest), interior cell (i,j) at (I,J) = (i+Np, j+Np).
//   - vx, vz, txx, tzz, txz hold the field value of every cell: the
//     interior value, or x+z of the split pair in a PML cell (the value
//     pml.c and propagate.c read as a neighbour).  Cells off the grid are 0.
//   - the ten split arrays (vx_x ... txz_z) are padded too and used only in
//     PML cells.
//   - b, bm, la, l2m, mu are padded: interior cells hold the staggered
//     averages of e_iso.c:438-462, PML cells the interior cell that pml.c
//     uses for them (the nearest interior cell, e.g. b[0], b[j], b[i*nz]),
//     and the lower-right corner's la/l2m optionally the cell nx*nx-1 that
//     pml.c:1044-1084 reads.
//   - damping: per padded index and stagger, the pair c1 = 2dt/(2+dt d),
//     c2 = (2-dt d)/(2dt) of pml.c:59-60, and whether the axis is inside
//     the PML there.  Along x, integer-stagger components (vx_x, txz_x) use
//     d[I] on the left and d[Np-1-(I-Np-nx)] on the right; half-stagger
//     components (vz_x, txx_x, tzz_x) use dh[I] and dH[Np-1-(I-Np-nx)];
//     same along z with d/dh/dH for (vx_z, txx_z, tzz_z) and (vz_z, txz_z).
#include <cmath>
#include <cstring>
#include <vector>

#include "e2d_fused.hpp"
#include "sm_internal.hpp"

namespace sm {

template <typename T>
struct E2dDev {
    int nx, nz, np, NX, NZ;
    T dt, h;
    T *vx, *vz, *txx, *tzz, *txz;
    T *vx_x, *vx_z, *vz_x, *vz_z, *txx_x, *txx_z, *tzz_x, *tzz_z, *txz_x, *txz_z;
    const T *b, *bm, *la, *l2m, *mu;
    // damping pairs: [c1, c2] interleaved, per padded index
    const T *dxi, *dxh, *dzi, *dzh;
};

template <typename T>
__device__ inline T at(const T *f, int I, int J, int NX, int NZ)
{
    return (I < 0 || J < 0 || I >= NX || J >= NZ) ? T(0) : f[(size_t)I * NZ + J];
}

// pml.c:59-63: u = ((u*c2) + coef*(t1-t2)/h)*c1 ; pml.c:123: u += dt*coef*(t1-t2)/h
template <typename T>
__device__ inline T split_step(T u, bool damped, const T *pair, T coef, T t1, T t2, T dt, T h)
{
    if (damped) {
        u *= pair[1];
        u += coef * (t1 - t2) / h;
        u *= pair[0];
        return u;
    }
    return u + dt * coef * (t1 - t2) / h;
}

template <typename T>
__global__ void __launch_bounds__(256) e2d_velocity(E2dDev<T> d)
{
    const int J = blockIdx.x * blockDim.x + threadIdx.x;
    const int I = blockIdx.y * blockDim.y + threadIdx.y;
    if (I >= d.NX || J >= d.NZ) return;
    const size_t c = (size_t)I * d.NZ + J;
    const bool px = I < d.np || I >= d.np + d.nx;
    const bool pz = J < d.np || J >= d.np + d.nz;
    const T h = d.h, dt = d.dt;
    if (!px && !pz) {
        // update_vx (propagate.c:2544-2547), update_vz (2562-2565)
        T dt1 = (d.txx[c] - at(d.txx, I - 1, J, d.NX, d.NZ)) / h;
        T dt2 = (d.txz[c] - at(d.txz, I, J - 1, d.NX, d.NZ)) / h;
        d.vx[c] += dt * d.b[c] * (dt1 + dt2);
        dt1 = (at(d.txz, I + 1, J, d.NX, d.NZ) - d.txz[c]) / h;
        dt2 = (at(d.tzz, I, J + 1, d.NX, d.NZ) - d.tzz[c]) / h;
        d.vz[c] += dt * d.bm[c] * (dt1 + dt2);
        return;
    }
    // pml_v (pml.c:39-680)
    const T b = d.b[c], bm = d.bm[c];
    T vxx = split_step(d.vx_x[c], px, d.dxi + 2 * I, b, d.txx[c], at(d.txx, I - 1, J, d.NX, d.NZ), dt, h);
    T vxz = split_step(d.vx_z[c], pz, d.dzi + 2 * J, b, d.txz[c], at(d.txz, I, J - 1, d.NX, d.NZ), dt, h);
    T vzx = split_step(d.vz_x[c], px, d.dxh + 2 * I, bm, at(d.txz, I + 1, J, d.NX, d.NZ), d.txz[c], dt, h);
    T vzz = split_step(d.vz_z[c], pz, d.dzh + 2 * J, bm, at(d.tzz, I, J + 1, d.NX, d.NZ), d.tzz[c], dt, h);
    d.vx_x[c] = vxx; d.vx_z[c] = vxz; d.vz_x[c] = vzx; d.vz_z[c] = vzz;
    d.vx[c] = vxx + vxz;
    d.vz[c] = vzx + vzz;
}

template <typename T>
__global__ void __launch_bounds__(256) e2d_stress(E2dDev<T> d)
{
    const int J = blockIdx.x * blockDim.x + threadIdx.x;
    const int I = blockIdx.y * blockDim.y + threadIdx.y;
    if (I >= d.NX || J >= d.NZ) return;
    const size_t c = (size_t)I * d.NZ + J;
    const bool px = I < d.np || I >= d.np + d.nx;
    const bool pz = J < d.np || J >= d.np + d.nz;
    const T h = d.h, dt = d.dt;
    const T vx1 = at(d.vx, I + 1, J, d.NX, d.NZ), vx0 = d.vx[c];
    const T vz0 = d.vz[c], vzm = at(d.vz, I, J - 1, d.NX, d.NZ);
    if (!px && !pz) {
        // update_txxzz (propagate.c:2606-2610), update_txz (2643-2646)
        T dvx_x = dt * (vx1 - vx0) / h;
        T dvz_z = dt * (vz0 - vzm) / h;
        d.txx[c] += d.l2m[c] * dvx_x + d.la[c] * dvz_z;
        d.tzz[c] += d.l2m[c] * dvz_z + d.la[c] * dvx_x;
        T dvx_z = dt * (at(d.vx, I, J + 1, d.NX, d.NZ) - vx0) / h;
        T dvz_x = dt * (vz0 - at(d.vz, I - 1, J, d.NX, d.NZ)) / h;
        d.txz[c] += d.mu[c] * (dvx_z + dvz_x);
        return;
    }
    // pml_t (pml.c:684-1403)
    const T la = d.la[c], l2m = d.l2m[c], mu = d.mu[c];
    T txxx = split_step(d.txx_x[c], px, d.dxh + 2 * I, l2m, vx1, vx0, dt, h);
    T tzzx = split_step(d.tzz_x[c], px, d.dxh + 2 * I, la, vx1, vx0, dt, h);
    T txxz = split_step(d.txx_z[c], pz, d.dzi + 2 * J, la, vz0, vzm, dt, h);
    T tzzz = split_step(d.tzz_z[c], pz, d.dzi + 2 * J, l2m, vz0, vzm, dt, h);
    T txzx = split_step(d.txz_x[c], px, d.dxi + 2 * I, mu, vz0, at(d.vz, I - 1, J, d.NX, d.NZ), dt, h);
    T txzz = split_step(d.txz_z[c], pz, d.dzh + 2 * J, mu, at(d.vx, I, J + 1, d.NX, d.NZ), vx0, dt, h);
    d.txx_x[c] = txxx; d.tzz_x[c] = tzzx; d.txx_z[c] = txxz; d.tzz_z[c] = tzzz;
    d.txz_x[c] = txzx; d.txz_z[c] = txzz;
    d.txx[c] = txxx + txxz;
    d.tzz[c] = tzzx + tzzz;
    d.txz[c] = txzx + txzz;
}

// add_src (src.c:170-184) then receivers, one block
template <typename T>
__global__ void __launch_bounds__(256)
e2d_src_rcv(E2dDev<T> d, int nsrc, const int64_t *sidx, const int64_t *soff, const int64_t *slen,
            const int *stype, const T *sfct, int nrec, const int64_t *ridx, const int *rcomp,
            T *trace, int64_t it, int64_t row)
{
    if (threadIdx.x == 0) {
        for (int n = 0; n < nsrc; ++n) {
            if (it >= slen[n]) continue;
            const T inc = d.dt * sfct[soff[n] + it];
            const int ty = stype[n];
            const size_t c = (size_t)sidx[n];
            if (ty == SM_SX || ty == SM_BULK) d.txx[c] += inc;
            if (ty == SM_SZ || ty == SM_BULK) d.tzz[c] += inc;
            if (ty == SM_SXZ) d.txz[c] += inc;
        }
    }
    __syncthreads();
    if (row < 0) return;
    for (int r = threadIdx.x; r < nrec; r += blockDim.x) {
        const int cp = rcomp[r];
        const T *f = cp == SM_VX ? d.vx : cp == SM_VZ ? d.vz : cp == SM_TXX ? d.txx
                   : cp == SM_TZZ ? d.tzz : d.txz;
        trace[row * nrec + r] = f[ridx[r]];
    }
}

}  // namespace sm

using namespace sm;

// Two device layouts:
//  - reference-order kernels (FP64, or FP32 with SM_E2D_NAIVE): fields = 5
//    totals + 10 split arrays, mat = b, bm, la, l2m, mu, damp = (c1, c2)
//    pairs, src/rcv as int64 records;
//  - fused FP32 kernel (default for FP32): fields = two parity buffers of
//    E2_NSTATE arrays + E2_NSPLIT stress split arrays, mat = the five
//    coefficients times dt/h, damp = (A, B) pairs, src/rcv as int4 records
//    sorted by padded cell.
struct sm_e2d {
    sm_ctx *ctx = nullptr;
    sm_e2d_params p{};
    int NX = 0, NZ = 0;
    size_t N = 0, elt = 0;
    bool fused = false;
    sm_dbuf fields;
    sm_dbuf mat;
    sm_dbuf damp;
    sm_dbuf src, rcv, trace;
    size_t step = 0, nsrc = 0, nrec = 0, max_steps = 0, rcv_step0 = 0;
    int chunk = 8, graph_steps = 32;
    int variant = SM_E2D_ROWS;
    sm::E2dRunner *runner = nullptr;
};

namespace {

void release(sm_e2d *s)
{
    dfree(s->fields); dfree(s->mat); dfree(s->damp); dfree(s->src); dfree(s->rcv); dfree(s->trace);
    delete s->runner;
    s->runner = nullptr;
}

// rows per chunk for a grid: enough waves to cover the chip about twice
int default_chunk(int NX, int NZ)
{
    const int ntz = (NZ + E2D_OWN - 1) / E2D_OWN;
    const int target = 2 * 256 * 4;
    int c = (int)((long long)NX * ntz / target);
    return std::max(2, std::min(16, c));
}

template <typename T>
E2dDev<T> view(sm_e2d *s)
{
    E2dDev<T> d;
    d.nx = (int)s->p.nx; d.nz = (int)s->p.nz; d.np = (int)s->p.npml;
    d.NX = s->NX; d.NZ = s->NZ;
    d.dt = (T)s->p.dt; d.h = (T)s->p.h;
    T *f = (T *)s->fields.ptr;
    const size_t N = s->N;
    T **dst[15] = {&d.vx, &d.vz, &d.txx, &d.tzz, &d.txz, &d.vx_x, &d.vx_z, &d.vz_x, &d.vz_z,
                   &d.txx_x, &d.txx_z, &d.tzz_x, &d.tzz_z, &d.txz_x, &d.txz_z};
    for (int k = 0; k < 15; ++k) *dst[k] = f + k * N;
    const T *m = (const T *)s->mat.ptr;
    d.b = m; d.bm = m + N; d.la = m + 2 * N; d.l2m = m + 3 * N; d.mu = m + 4 * N;
    const T *q = (const T *)s->damp.ptr;
    d.dxi = q; d.dxh = q + 2 * s->NX; d.dzi = q + 4 * s->NX; d.dzh = q + 4 * s->NX + 2 * s->NZ;
    return d;
}

template <typename T>
int upload(void *dst, const std::vector<double> &src, hipStream_t st)
{
    std::vector<T> tmp(src.size());
    for (size_t i = 0; i < src.size(); ++i) tmp[i] = (T)src[i];
    SM_HIP(hipMemcpyAsync(dst, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice, st));
    SM_HIP(hipStreamSynchronize(st));
    return SM_OK;
}

template <typename T>
int run(sm_e2d *s, size_t nsteps)
{
    hipStream_t st = s->ctx->stream;
    E2dDev<T> d = view<T>(s);
    const dim3 blk(64, 4, 1);
    const dim3 grd((s->NZ + 63) / 64, (s->NX + 3) / 4, 1);
    const int64_t *sp = (const int64_t *)s->src.ptr;
    const int *stype = (const int *)(sp + 3 * s->nsrc);
    const T *sfct = (const T *)(sp + 4 * s->nsrc);
    const int64_t *ridx = (const int64_t *)s->rcv.ptr;
    const int *rcomp = (const int *)(ridx + s->nrec);
    for (size_t n = 0; n < nsteps; ++n) {
        const size_t it = s->step + n;
        e2d_velocity<T><<<grd, blk, 0, st>>>(d);
        e2d_stress<T><<<grd, blk, 0, st>>>(d);
        int64_t row = -1;
        if (s->nrec && it >= s->rcv_step0 && it - s->rcv_step0 < s->max_steps)
            row = (int64_t)(it - s->rcv_step0);
        if (s->nsrc || row >= 0)
            e2d_src_rcv<T><<<1, 256, 0, st>>>(d, (int)s->nsrc, sp, sp + s->nsrc, sp + 2 * s->nsrc,
                                              stype, sfct, (int)s->nrec, ridx, rcomp,
                                              (T *)s->trace.ptr, (int64_t)it, row);
    }
    SM_HIP(hipGetLastError());
    s->step += nsteps;
    return SM_OK;
}

}  // namespace

namespace sm {
// Descriptor of a fused-layout solver for the next nsteps steps.
E2dShot e2d_shot(sm_e2d *s, size_t nsteps)
{
    E2dShot d;
    std::memset(&d, 0, sizeof(d));
    d.NX = s->NX; d.NZ = s->NZ;
    d.np = (int)s->p.npml; d.nx = (int)s->p.nx; d.nz = (int)s->p.nz;
    d.chunk = std::max(1, std::min(s->chunk, s->NX));
    e2d_items(d.NX, d.NZ, d.chunk, &d.ntz, &d.nch);
    d.nsrc = (int)s->nsrc;
    d.nrec = (int)s->nrec;
    d.it0 = (int)s->step;
    d.nsteps = (int)nsteps;
    d.rcv0 = (int)s->rcv_step0;
    d.rmax = (int)s->max_steps;
    d.N = (unsigned)s->N;
    float *f = (float *)s->fields.ptr;
    d.st[0] = f;
    d.st[1] = f + E2_NSTATE * s->N;
    d.ss = f + 2 * E2_NSTATE * s->N;
    d.mat = (const float *)s->mat.ptr;
    d.dmp = (const float *)s->damp.ptr;
    d.src = (const int4 *)s->src.ptr;
    d.srow = s->src.ptr ? (const int *)((const int4 *)s->src.ptr + s->nsrc) : nullptr;
    d.sfct = s->src.ptr ? (const float *)(d.srow + s->NX + 1) : nullptr;
    d.rcv = (const int4 *)s->rcv.ptr;
    d.rrow = s->rcv.ptr ? (const int *)((const int4 *)s->rcv.ptr + s->nrec) : nullptr;
    d.trace = (float *)s->trace.ptr;
    return d;
}

int e2d_check_steps(const sm_e2d *s, size_t nsteps)
{
    SM_CHECK(s->step + nsteps < (1ull << 31), SM_ERR_ARG,
             "step count %zu beyond the fused kernel's 2^31 limit", s->step + nsteps);
    return SM_OK;
}
}  // namespace sm

namespace {

int run_fused(sm_e2d *s, size_t nsteps)
{
    int rc = e2d_check_steps(s, nsteps);
    if (rc != SM_OK) return rc;
    if (!s->runner) s->runner = new E2dRunner;
    std::vector<E2dShot> one{e2d_shot(s, nsteps)};
    s->runner->staged = s->variant == SM_E2D_STAGED;
    rc = s->runner->run(one, s->ctx->stream, s->graph_steps);
    if (rc != SM_OK) return rc;
    s->step += nsteps;
    return SM_OK;
}

template <typename T>
int read_back(sm_e2d *s, const void *dev, size_t n, double *out)
{
    std::vector<T> tmp(n);
    SM_HIP(hipMemcpyAsync(tmp.data(), dev, n * sizeof(T), hipMemcpyDeviceToHost, s->ctx->stream));
    SM_HIP(hipStreamSynchronize(s->ctx->stream));
    for (size_t i = 0; i < n; ++i) out[i] = (double)tmp[i];
    return SM_OK;
}

int zero_state(sm_e2d *s)
{
    hipStream_t st = s->ctx->stream;
    SM_HIP(hipMemsetAsync(s->fields.ptr, 0, s->fields.bytes, st));
    if (s->trace.ptr) SM_HIP(hipMemsetAsync(s->trace.ptr, 0, s->trace.bytes, st));
    SM_HIP(hipStreamSynchronize(st));
    s->step = 0;
    s->rcv_step0 = 0;
    return SM_OK;
}

int comp_slot(int comp)
{
    switch (comp) {
    case SM_VX: return 0;
    case SM_VZ: return 1;
    case SM_TXX: return 2;
    case SM_TZZ: return 3;
    case SM_TXZ: return 4;
    default: return -1;
    }
}

}  // namespace

extern "C" {

// The fused kernel masks an access by passing the per-lane offset 2^31
// (e2d_fused.hip): that is out of range only while the largest buffer
// resource, the two state parities plus the split fields, stays below
// 2^31 bytes.
static bool e2d_fused_fits(size_t N)
{
    return (2 * E2_NSTATE + E2_NSPLIT) * N * 4 < (1ull << 31);
}

size_t sm_e2d_bytes_needed(const sm_e2d_params *p, size_t nrec, size_t max_steps)
{
    if (!p) return 0;
    const size_t e = p->storage == SM_STORE_F64 ? 8 : 4;
    const size_t NX = p->nx + 2 * p->npml, NZ = p->nz + 2 * p->npml, N = NX * NZ;
    if (p->storage == SM_STORE_F32 && !(p->flags & SM_E2D_NAIVE)) {
        if (!e2d_fused_fits(N)) return 0;   // sm_e2d_create refuses the grid
        return 4 * ((2 * E2_NSTATE + E2_NSPLIT + E2_NMAT) * N + 4 * (NX + NZ) + nrec * max_steps) +
               16 * nrec + 4096;
    }
    return e * (20 * N + 4 * (NX + NZ) + nrec * max_steps) + 16 * nrec;
}

int sm_e2d_create(sm_ctx *ctx, const sm_e2d_params *pp, const double *vp, const double *vs,
                  const double *rho, sm_e2d **out)
{
    SM_CHECK(ctx && pp && vp && vs && rho && out, SM_ERR_ARG, "sm_e2d_create: NULL argument");
    const sm_e2d_params &p = *pp;
    SM_CHECK(p.nx >= 2 && p.nz >= 2, SM_ERR_ARG, "sm_e2d_create: grid must be at least 2x2");
    SM_CHECK(p.npml >= 1, SM_ERR_ARG,
             "sm_e2d_create: e_iso needs at least one PML layer (abs 1; Npml-1 underflows otherwise)");
    SM_CHECK(p.h > 0 && p.dt > 0, SM_ERR_ARG, "sm_e2d_create: h and dt must be > 0");
    SM_CHECK(p.storage == SM_STORE_F32 || p.storage == SM_STORE_F64, SM_ERR_ARG,
             "sm_e2d_create: storage %d not supported", p.storage);
    SM_CHECK((p.flags & ~SM_E2D_NAIVE) == 0, SM_ERR_ARG, "sm_e2d_create: unknown flags %#x", p.flags);
    const size_t nx = p.nx, nz = p.nz, np = p.npml, n = nx * nz;
    const size_t NX = nx + 2 * np, NZ = nz + 2 * np, N = NX * NZ;
    SM_CHECK(NX < (1u << 20) && NZ < (1u << 20), SM_ERR_ARG, "sm_e2d_create: grid too large");
    for (size_t c = 0; c < n; ++c)
        SM_CHECK(rho[c] > 0.0, SM_ERR_ARG, "sm_e2d_create: rho must be > 0 (cell %zu)", c);

    // Lame constants and staggered averages, e_iso.c:423-462
    std::vector<double> la(n), mu(n), l2m(n), b(n), bm(n), lt(n), mt(n), l2t(n);
    double vpmax = 0.0;
    for (size_t c = 0; c < n; ++c) {
        vpmax = vp[c] > vpmax ? vp[c] : vpmax;
        lt[c] = rho[c] * (vp[c] * vp[c] - 2. * vs[c] * vs[c]);
        mt[c] = rho[c] * vs[c] * vs[c];
        l2t[c] = lt[c] + 2. * mt[c];
        b[c] = 1. / rho[c];
    }
    for (size_t i = 0; i < nx; ++i) {
        for (size_t j = 0; j + 1 < nz; ++j) mu[i * nz + j] = (mt[i * nz + j] + mt[i * nz + j + 1]) / 2.0;
        mu[i * nz + nz - 1] = mt[i * nz + nz - 1];
    }
    for (size_t j = 0; j < nz; ++j) {
        for (size_t i = 0; i + 1 < nx; ++i) {
            la[i * nz + j] = (lt[i * nz + j] + lt[(i + 1) * nz + j]) / 2.0;
            l2m[i * nz + j] = (l2t[i * nz + j] + l2t[(i + 1) * nz + j]) / 2.0;
        }
        la[(nx - 1) * nz + j] = lt[(nx - 1) * nz + j];
        l2m[(nx - 1) * nz + j] = l2t[(nx - 1) * nz + j];
    }
    for (size_t i = 0; i + 1 < nx; ++i) {
        for (size_t j = 0; j + 1 < nz; ++j)
            bm[i * nz + j] = (b[i * nz + j] + b[(i + 1) * nz + j] + b[i * nz + j + 1] +
                              b[(i + 1) * nz + j + 1]) / 4.0;
        bm[i * nz + nz - 1] = (b[i * nz + nz - 1] + b[(i + 1) * nz + nz - 1]) / 2.0;
    }
    for (size_t j = 0; j + 1 < nz; ++j)
        bm[(nx - 1) * nz + j] = (b[(nx - 1) * nz + j] + b[(nx - 1) * nz + j + 1]) / 2.0;
    bm[n - 1] = b[n - 1];

    // CFL, e_iso.c:476-479
    if (p.dt > p.h / (std::sqrt(2.0) * vpmax))
        return fail(SM_ERR_CFL, "time step too large: dt %g > h/(sqrt(2) vpmax) = %g", p.dt,
                    p.h / (std::sqrt(2.0) * vpmax));

    // damping from the mean Vp, e_iso.c:486-500
    double vmean = 0.0;
    for (size_t c = 0; c < n; ++c) vmean += vp[c];
    vmean /= (double)nx * nz;
    std::vector<double> d(np), dh(np), dH(np + 1);
    {
        const double R = 0.001;
        const int pow_d = 2;
        const double delta = np * p.h;
        const double d0 = std::log(1. / R) * 3. / 2. * vmean / delta;
        for (size_t k = 0; k < np; ++k) {
            d[k] = d0 * std::pow(((np - k) * p.h / delta), pow_d);
            dh[k] = d0 * std::pow((((np - k) * p.h - p.h / 2.) / delta), pow_d);
        }
        for (size_t k = 0; k < np + 1; ++k)
            dH[k] = d0 * std::pow((((np - k) * p.h + p.h / 2.) / delta), pow_d);
        if (p.damping_scale != 1.0) {
            for (auto &v : d) v *= p.damping_scale;
            for (auto &v : dh) v *= p.damping_scale;
            for (auto &v : dH) v *= p.damping_scale;
        }
    }
    // per padded index (c1, c2) pairs; undamped positions keep (0, 0)
    std::vector<double> dt_tab(4 * NX + 4 * NZ, 0.0);
    auto pair = [&](double *dst, double dd) {
        dst[0] = (2. * p.dt) / (2. + p.dt * dd);
        dst[1] = (2. - p.dt * dd) / (2. * p.dt);
    };
    {
        double *xi = dt_tab.data(), *xh = xi + 2 * NX, *zi = xi + 4 * NX, *zh = zi + 2 * NZ;
        for (size_t I = 0; I < NX; ++I) {
            if (I < np) { pair(xi + 2 * I, d[I]); pair(xh + 2 * I, dh[I]); }
            else if (I >= np + nx) {
                const size_t r = np - 1 - (I - np - nx);
                pair(xi + 2 * I, d[r]); pair(xh + 2 * I, dH[r]);
            }
        }
        for (size_t J = 0; J < NZ; ++J) {
            if (J < np) { pair(zi + 2 * J, d[J]); pair(zh + 2 * J, dh[J]); }
            else if (J >= np + nz) {
                const size_t r = np - 1 - (J - np - nz);
                pair(zi + 2 * J, d[r]); pair(zh + 2 * J, dH[r]);
            }
        }
    }
    // padded material: PML cells take the nearest interior cell
    std::vector<double> mat(5 * N);
    const size_t q1 = (p.q1_compat && nx * nx <= n) ? nx * nx - 1 : n - 1;
    for (size_t I = 0; I < NX; ++I)
        for (size_t J = 0; J < NZ; ++J) {
            const size_t ci = I < np ? 0 : (I >= np + nx ? nx - 1 : I - np);
            const size_t cj = J < np ? 0 : (J >= np + nz ? nz - 1 : J - np);
            size_t m = ci * nz + cj;
            const size_t c = I * NZ + J;
            mat[c] = b[m];
            mat[N + c] = bm[m];
            const bool lr = I >= np + nx && J >= np + nz;   // pml.c:1037-1087
            const size_t mq = lr ? q1 : m;
            mat[2 * N + c] = la[mq];
            mat[3 * N + c] = l2m[mq];
            mat[4 * N + c] = mu[m];
        }

    SM_CHECK(set_device(ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    sm_e2d *s = new sm_e2d;
    s->ctx = ctx;
    s->p = p;
    s->NX = (int)NX; s->NZ = (int)NZ; s->N = N;
    s->elt = p.storage == SM_STORE_F64 ? 8 : 4;
    s->fused = p.storage == SM_STORE_F32 && !(p.flags & SM_E2D_NAIVE);
    s->chunk = default_chunk((int)NX, (int)NZ);
    int rc = SM_OK;
    if (s->fused) {
        // coefficients times dt/h (the interior forms multiply by dt/h,
        // the damped split forms by c1 = 2dt/(2+dt d) and 1/h)
        const double f = p.dt / p.h;
        for (auto &v : mat) v *= f;
        // (A, B) = (c1*c2, c1/dt) where damped, (1, 1) where not
        std::vector<double> ab(4 * NX + 4 * NZ, 1.0);
        auto put = [&](double *dst, double dd) {
            dst[0] = (2. - p.dt * dd) / (2. + p.dt * dd);
            dst[1] = 2. / (2. + p.dt * dd);
        };
        double *xi = ab.data(), *xh = xi + 2 * NX, *zi = xi + 4 * NX, *zh = zi + 2 * NZ;
        for (size_t I = 0; I < NX; ++I) {
            if (I < np) { put(xi + 2 * I, d[I]); put(xh + 2 * I, dh[I]); }
            else if (I >= np + nx) {
                const size_t r = np - 1 - (I - np - nx);
                put(xi + 2 * I, d[r]); put(xh + 2 * I, dH[r]);
            }
        }
        for (size_t J = 0; J < NZ; ++J) {
            if (J < np) { put(zi + 2 * J, d[J]); put(zh + 2 * J, dh[J]); }
            else if (J >= np + nz) {
                const size_t r = np - 1 - (J - np - nz);
                put(zi + 2 * J, d[r]); put(zh + 2 * J, dH[r]);
            }
        }
        SM_CHECK(e2d_fused_fits(N), SM_ERR_ARG,
                 "sm_e2d_create: padded grid %zux%zu too large for the fused FP32 kernel "
                 "(state and split fields must stay below 2 GiB; use SM_E2D_NAIVE or FP64)", NX, NZ);
        rc = dalloc(s->fields, (2 * E2_NSTATE + E2_NSPLIT) * N * 4);
        // 16 floats past the last material array: the staged kernel's
        // 16-byte chunks of the last row may extend into them
        if (rc == SM_OK) rc = dalloc(s->mat, (E2_NMAT * N + 16) * 4);
        if (rc == SM_OK) rc = dalloc(s->damp, ab.size() * 4);
        if (rc == SM_OK) rc = upload<float>(s->mat.ptr, mat, ctx->stream);
        if (rc == SM_OK) rc = upload<float>(s->damp.ptr, ab, ctx->stream);
    } else {
        rc = dalloc(s->fields, 15 * N * s->elt);
        if (rc == SM_OK) rc = dalloc(s->mat, 5 * N * s->elt);
        if (rc == SM_OK) rc = dalloc(s->damp, dt_tab.size() * s->elt);
        if (rc == SM_OK) {
            if (s->elt == 8) {
                rc = upload<double>(s->mat.ptr, mat, ctx->stream);
                if (rc == SM_OK) rc = upload<double>(s->damp.ptr, dt_tab, ctx->stream);
            } else {
                rc = upload<float>(s->mat.ptr, mat, ctx->stream);
                if (rc == SM_OK) rc = upload<float>(s->damp.ptr, dt_tab, ctx->stream);
            }
        }
    }
    if (rc == SM_OK) rc = zero_state(s);
    if (rc != SM_OK) { release(s); delete s; return rc; }
    *out = s;
    return SM_OK;
}

int sm_e2d_destroy(sm_e2d *s)
{
    if (!s) return SM_OK;
    (void)set_device(s->ctx);
    (void)hipStreamSynchronize(s->ctx->stream);
    release(s);
    delete s;
    return SM_OK;
}

int sm_e2d_set_sources(sm_e2d *s, size_t nsrc, const int *type, const size_t *idx,
                       const size_t *length, const double *fct)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_e2d_set_sources: NULL solver");
    SM_CHECK(nsrc == 0 || (type && idx && length && fct), SM_ERR_ARG,
             "sm_e2d_set_sources: NULL array");
    SM_CHECK(nsrc < (1u << 20), SM_ERR_ARG, "sm_e2d_set_sources: too many sources");
    const size_t n = s->p.nx * s->p.nz;
    size_t total = 0;
    for (size_t k = 0; k < nsrc; ++k) {
        SM_CHECK(idx[k] < n, SM_ERR_ARG, "source %zu: index %zu outside the %zux%zu grid", k, idx[k],
                 s->p.nx, s->p.nz);
        // enum sourceType (structs.h:44) has 13 values; add_src acts on
        // four of them and ignores the others, values outside are errors
        SM_CHECK(type[k] >= 0 && type[k] <= 12, SM_ERR_ARG, "source %zu: type %d is not a source type",
                 k, type[k]);
        total += length[k];
    }
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    SM_HIP(hipStreamSynchronize(s->ctx->stream));
    dfree(s->src);
    s->nsrc = 0;
    if (nsrc == 0) return SM_OK;
    if (s->fused) {
        // int4 (padded cell, type, length, offset) sorted by cell, then dt*fct
        SM_CHECK(total < (1ull << 31), SM_ERR_ARG, "sm_e2d_set_sources: source samples too many");
        std::vector<int> order(nsrc);
        for (size_t k = 0; k < nsrc; ++k) order[k] = (int)k;
        auto cell = [&](size_t k) {
            const size_t i = idx[k] / s->p.nz, j = idx[k] % s->p.nz;
            return (int)((i + s->p.npml) * s->NZ + j + s->p.npml);
        };
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cell(a) < cell(b); });
        std::vector<size_t> offs(nsrc);
        for (size_t k = 0, o = 0; k < nsrc; ++k) { offs[k] = o; o += length[k]; }
        const size_t NXp = (size_t)s->NX + 1;
        const size_t bytes = nsrc * 16 + NXp * 4 + total * 4;
        int rc = dalloc(s->src, bytes);
        if (rc != SM_OK) return rc;
        std::vector<unsigned char> hb(bytes);
        int *rec = (int *)hb.data();
        int *srow = (int *)(hb.data() + nsrc * 16);
        float *hf = (float *)(hb.data() + nsrc * 16 + NXp * 4);
        for (size_t q = 0; q < nsrc; ++q) {
            const int k = order[q];
            rec[4 * q] = cell(k);
            rec[4 * q + 1] = type[k];
            rec[4 * q + 2] = (int)std::min<size_t>(length[k], 1u << 30);
            rec[4 * q + 3] = (int)offs[k];
        }
        for (size_t I = 0, q = 0; I < NXp; ++I) {
            while (q < nsrc && rec[4 * q] < (int)(I * s->NZ)) ++q;
            srow[I] = (int)q;
        }
        for (size_t q = 0; q < total; ++q) hf[q] = (float)(s->p.dt * fct[q]);
        SM_HIP(hipMemcpyAsync(s->src.ptr, hb.data(), bytes, hipMemcpyHostToDevice, s->ctx->stream));
        SM_HIP(hipStreamSynchronize(s->ctx->stream));
        s->nsrc = nsrc;
        return SM_OK;
    }
    const size_t bytes = 4 * nsrc * sizeof(int64_t) + total * s->elt;
    int rc = dalloc(s->src, bytes);
    if (rc != SM_OK) return rc;
    std::vector<unsigned char> hb(bytes);
    int64_t *hi = (int64_t *)hb.data();
    int *ht = (int *)(hi + 3 * nsrc);
    size_t off = 0;
    for (size_t k = 0; k < nsrc; ++k) {
        const size_t i = idx[k] / s->p.nz, j = idx[k] % s->p.nz;
        hi[k] = (int64_t)((i + s->p.npml) * s->NZ + j + s->p.npml);
        hi[nsrc + k] = (int64_t)off;
        hi[2 * nsrc + k] = (int64_t)length[k];
        ht[k] = type[k];
        off += length[k];
    }
    for (size_t q = 0; q < total; ++q) {
        if (s->elt == 8) ((double *)(hi + 4 * nsrc))[q] = fct[q];
        else ((float *)(hi + 4 * nsrc))[q] = (float)fct[q];
    }
    SM_HIP(hipMemcpyAsync(s->src.ptr, hb.data(), bytes, hipMemcpyHostToDevice, s->ctx->stream));
    SM_HIP(hipStreamSynchronize(s->ctx->stream));
    s->nsrc = nsrc;
    return SM_OK;
}

int sm_e2d_set_receivers(sm_e2d *s, size_t nrec, const size_t *idx, const int *comp,
                         size_t max_steps)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_e2d_set_receivers: NULL solver");
    SM_CHECK(nrec == 0 || (idx && comp), SM_ERR_ARG, "sm_e2d_set_receivers: NULL array");
    SM_CHECK(nrec < (1u << 24), SM_ERR_ARG, "sm_e2d_set_receivers: too many receivers");
    const size_t n = s->p.nx * s->p.nz;
    for (size_t r = 0; r < nrec; ++r) {
        SM_CHECK(idx[r] < n, SM_ERR_ARG, "receiver %zu: index %zu outside the grid", r, idx[r]);
        SM_CHECK(comp_slot(comp[r]) >= 0, SM_ERR_ARG, "receiver %zu: component %d not recordable in 2D",
                 r, comp[r]);
    }
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    SM_HIP(hipStreamSynchronize(s->ctx->stream));
    dfree(s->rcv);
    dfree(s->trace);
    s->nrec = 0;
    s->max_steps = 0;
    if (nrec == 0 || max_steps == 0) return SM_OK;
    if (s->fused) {
        // int4 (padded cell, slot, trace column, 0) sorted by cell
        SM_CHECK(max_steps < (1ull << 31), SM_ERR_ARG, "sm_e2d_set_receivers: max_steps too large");
        std::vector<int> hr(4 * nrec + s->NX + 1);
        std::vector<int> order(nrec);
        for (size_t r = 0; r < nrec; ++r) order[r] = (int)r;
        auto cell = [&](size_t r) {
            const size_t i = idx[r] / s->p.nz, j = idx[r] % s->p.nz;
            return (int)((i + s->p.npml) * s->NZ + j + s->p.npml);
        };
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cell(a) < cell(b); });
        for (size_t q = 0; q < nrec; ++q) {
            const int r = order[q];
            hr[4 * q] = cell(r);
            hr[4 * q + 1] = comp_slot(comp[r]);
            hr[4 * q + 2] = r;
            hr[4 * q + 3] = 0;
        }
        for (size_t I = 0, q = 0; I <= (size_t)s->NX; ++I) {
            while (q < nrec && hr[4 * q] < (int)(I * s->NZ)) ++q;
            hr[4 * nrec + I] = (int)q;
        }
        int rc = dalloc(s->rcv, hr.size() * 4);
        if (rc == SM_OK) rc = dalloc(s->trace, nrec * max_steps * 4);
        if (rc != SM_OK) { dfree(s->rcv); dfree(s->trace); return rc; }
        SM_HIP(hipMemcpyAsync(s->rcv.ptr, hr.data(), hr.size() * 4, hipMemcpyHostToDevice,
                              s->ctx->stream));
        SM_HIP(hipStreamSynchronize(s->ctx->stream));
        s->nrec = nrec;
        s->max_steps = max_steps;
        s->rcv_step0 = s->step;
        return SM_OK;
    }
    int rc = dalloc(s->rcv, nrec * (sizeof(int64_t) + sizeof(int)));
    if (rc == SM_OK) rc = dalloc(s->trace, nrec * max_steps * s->elt);
    if (rc != SM_OK) { dfree(s->rcv); dfree(s->trace); return rc; }
    std::vector<unsigned char> hb(nrec * (sizeof(int64_t) + sizeof(int)));
    int64_t *hi = (int64_t *)hb.data();
    int *hc = (int *)(hi + nrec);
    for (size_t r = 0; r < nrec; ++r) {
        const size_t i = idx[r] / s->p.nz, j = idx[r] % s->p.nz;
        hi[r] = (int64_t)((i + s->p.npml) * s->NZ + j + s->p.npml);
        hc[r] = comp[r];
    }
    SM_HIP(hipMemcpyAsync(s->rcv.ptr, hb.data(), hb.size(), hipMemcpyHostToDevice, s->ctx->stream));
    SM_HIP(hipMemsetAsync(s->trace.ptr, 0, s->trace.bytes, s->ctx->stream));
    SM_HIP(hipStreamSynchronize(s->ctx->stream));
    s->nrec = nrec;
    s->max_steps = max_steps;
    s->rcv_step0 = s->step;
    return SM_OK;
}

int sm_e2d_run(sm_e2d *s, size_t nsteps)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_e2d_run: NULL solver");
    if (nsteps == 0) return SM_OK;
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    if (s->fused) return run_fused(s, nsteps);
    if (s->elt == 8) return run<double>(s, nsteps);
    return run<float>(s, nsteps);
}

int sm_e2d_current_step(const sm_e2d *s, size_t *it)
{
    SM_CHECK(s && it, SM_ERR_ARG, "sm_e2d_current_step: NULL argument");
    *it = s->step;
    return SM_OK;
}

int sm_e2d_get_traces(sm_e2d *s, size_t it0, size_t count, double *out)
{
    SM_CHECK(s && (out || count == 0), SM_ERR_ARG, "sm_e2d_get_traces: NULL argument");
    SM_CHECK(it0 + count <= s->max_steps, SM_ERR_ARG,
             "sm_e2d_get_traces: steps [%zu,%zu) beyond the %zu recorded", it0, it0 + count,
             s->max_steps);
    if (count == 0 || s->nrec == 0) return SM_OK;
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    const char *src = (const char *)s->trace.ptr + it0 * s->nrec * s->elt;
    if (s->elt == 8) return read_back<double>(s, src, count * s->nrec, out);
    return read_back<float>(s, src, count * s->nrec, out);
}

int sm_e2d_get_field_padded(sm_e2d *s, int comp, double *out)
{
    SM_CHECK(s && out, SM_ERR_ARG, "sm_e2d_get_field_padded: NULL argument");
    const int slot = comp_slot(comp);
    SM_CHECK(slot >= 0, SM_ERR_ARG, "sm_e2d_get_field_padded: component %d", comp);
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    size_t arr = (size_t)slot;
    if (s->fused) arr += (s->step & 1) * E2_NSTATE;   // parity buffer holding step s->step
    const char *src = (const char *)s->fields.ptr + arr * s->N * s->elt;
    if (s->elt == 8) return read_back<double>(s, src, s->N, out);
    return read_back<float>(s, src, s->N, out);
}

int sm_e2d_get_field(sm_e2d *s, int comp, double *out)
{
    SM_CHECK(s && out, SM_ERR_ARG, "sm_e2d_get_field: NULL argument");
    std::vector<double> pad(s->N);
    int rc = sm_e2d_get_field_padded(s, comp, pad.data());
    if (rc != SM_OK) return rc;
    const size_t nx = s->p.nx, nz = s->p.nz, np = s->p.npml;
    for (size_t i = 0; i < nx; ++i)
        std::memcpy(out + i * nz, pad.data() + (i + np) * s->NZ + np, nz * sizeof(double));
    return SM_OK;
}

int sm_e2d_reset(sm_e2d *s)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_e2d_reset: NULL solver");
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    return zero_state(s);
}

int sm_e2d_device_bytes(const sm_e2d *s, size_t *bytes)
{
    SM_CHECK(s && bytes, SM_ERR_ARG, "sm_e2d_device_bytes: NULL argument");
    *bytes = s->fields.bytes + s->mat.bytes + s->damp.bytes + s->src.bytes + s->rcv.bytes +
             s->trace.bytes + (s->runner ? s->runner->device_bytes() : 0);
    return SM_OK;
}

int sm_e2d_set_launch(sm_e2d *s, int rows_per_chunk, int graph_steps)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_e2d_set_launch: NULL solver");
    SM_CHECK(s->fused, SM_ERR_ARG, "sm_e2d_set_launch: only the fused FP32 kernel has launch parameters");
    SM_CHECK(rows_per_chunk >= 1 && rows_per_chunk <= 4096, SM_ERR_ARG,
             "sm_e2d_set_launch: rows_per_chunk %d outside [1, 4096]", rows_per_chunk);
    SM_CHECK(graph_steps >= 0 && graph_steps <= 4096, SM_ERR_ARG,
             "sm_e2d_set_launch: graph_steps %d outside [0, 4096]", graph_steps);
    s->chunk = rows_per_chunk;
    s->graph_steps = graph_steps;
    return SM_OK;
}

int sm_e2d_set_variant(sm_e2d *s, int variant)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_e2d_set_variant: NULL solver");
    SM_CHECK(s->fused, SM_ERR_ARG, "sm_e2d_set_variant: only the fused FP32 kernel has variants");
    SM_CHECK(variant == SM_E2D_ROWS || variant == SM_E2D_STAGED, SM_ERR_ARG,
             "sm_e2d_set_variant: unknown variant %d", variant);
    s->variant = variant;
    return SM_OK;
}

int sm_e2d_get_variant(const sm_e2d *s, int *variant)
{
    SM_CHECK(s && variant, SM_ERR_ARG, "sm_e2d_get_variant: NULL argument");
    *variant = s->fused ? s->variant : -1;
    return SM_OK;
}

int sm_e2d_get_launch(const sm_e2d *s, int *rows_per_chunk, int *graph_steps)
{
    SM_CHECK(s && rows_per_chunk && graph_steps, SM_ERR_ARG, "sm_e2d_get_launch: NULL argument");
    *rows_per_chunk = s->chunk;
    *graph_steps = s->graph_steps;
    return SM_OK;
}

int sm_e2d_set_timing(sm_e2d *s, int on)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_e2d_set_timing: NULL solver");
    SM_CHECK(s->fused, SM_ERR_ARG, "sm_e2d_set_timing: only the fused FP32 kernel is timed");
    if (!s->runner) s->runner = new E2dRunner;
    s->runner->timing = on != 0;
    s->runner->ms_total = 0.0;
    s->runner->launches = 0;
    return SM_OK;
}

int sm_e2d_get_timing(const sm_e2d *s, double *ms_total, size_t *launches)
{
    SM_CHECK(s && ms_total && launches, SM_ERR_ARG, "sm_e2d_get_timing: NULL argument");
    *ms_total = s->runner ? s->runner->ms_total : 0.0;
    *launches = s->runner ? s->runner->launches : 0;
    return SM_OK;
}

}  // extern "C"

struct sm_e2d_batch {
    sm_ctx *ctx = nullptr;
    sm::E2dRunner runner;
    int graph_steps = 32;
    int variant = SM_E2D_ROWS;
    size_t group_bytes = 0;         // 0: all shots in every launch
};

extern "C" {

int sm_e2d_batch_create(sm_ctx *ctx, sm_e2d_batch **out)
{
    SM_CHECK(ctx && out, SM_ERR_ARG, "sm_e2d_batch_create: NULL argument");
    sm_e2d_batch *b = new sm_e2d_batch;
    b->ctx = ctx;
    *out = b;
    return SM_OK;
}

int sm_e2d_batch_destroy(sm_e2d_batch *b)
{
    if (!b) return SM_OK;
    (void)set_device(b->ctx);
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
    return SM_OK;
}

int sm_e2d_batch_set_group_bytes(sm_e2d_batch *b, size_t bytes)
{
    SM_CHECK(b, SM_ERR_ARG, "sm_e2d_batch_set_group_bytes: NULL batch");
    b->group_bytes = bytes;
    return SM_OK;
}

int sm_e2d_batch_set_graph_steps(sm_e2d_batch *b, int graph_steps)
{
    SM_CHECK(b, SM_ERR_ARG, "sm_e2d_batch_set_graph_steps: NULL batch");
    SM_CHECK(graph_steps >= 0 && graph_steps <= 4096, SM_ERR_ARG,
             "sm_e2d_batch_set_graph_steps: %d outside [0, 4096]", graph_steps);
    b->graph_steps = graph_steps;
    return SM_OK;
}

// Checks the shots of a batch run and builds their descriptors; work
// queued on the shots' own streams is complete when it returns.
static int batch_shots(sm_e2d_batch *b, size_t nshots, sm_e2d *const *shots, const size_t *nsteps,
                       std::vector<E2dShot> &d)
{
    SM_CHECK(b && (nshots == 0 || (shots && nsteps)), SM_ERR_ARG, "sm_e2d_batch_run: NULL argument");
    d.reserve(nshots);
    for (size_t k = 0; k < nshots; ++k) {
        sm_e2d *s = shots[k];
        SM_CHECK(s, SM_ERR_ARG, "sm_e2d_batch_run: shot %zu is NULL", k);
        SM_CHECK(s->fused, SM_ERR_ARG, "sm_e2d_batch_run: shot %zu is not an FP32 fused solver", k);
        SM_CHECK(s->ctx->device == b->ctx->device, SM_ERR_ARG,
                 "sm_e2d_batch_run: shot %zu is on device %d, the batch on %d", k, s->ctx->device,
                 b->ctx->device);
        for (size_t q = 0; q < k; ++q)
            SM_CHECK(shots[q] != s, SM_ERR_ARG, "sm_e2d_batch_run: shot %zu listed twice", k);
        int rc = e2d_check_steps(s, nsteps[k]);
        if (rc != SM_OK) return rc;
        d.push_back(e2d_shot(s, nsteps[k]));
    }
    SM_CHECK(set_device(b->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    // work queued on the shots' own streams completes before the batch reads them
    for (size_t k = 0; k < nshots; ++k)
        if (shots[k]->ctx->stream != b->ctx->stream) SM_HIP(hipStreamSynchronize(shots[k]->ctx->stream));
    b->runner.staged = b->variant == SM_E2D_STAGED;
    return SM_OK;
}

// Device memory a shot's step kernels touch: wavefields, material, damping.
static size_t shot_footprint(const sm_e2d *s)
{
    return s->fields.bytes + s->mat.bytes + s->damp.bytes;
}

int sm_e2d_batch_run(sm_e2d_batch *b, size_t nshots, sm_e2d *const *shots, const size_t *nsteps)
{
    std::vector<E2dShot> d;
    int rc = batch_shots(b, nshots, shots, nsteps, d);
    if (rc != SM_OK) return rc;
    if (b->group_bytes == 0) {
        rc = b->runner.run(d, b->ctx->stream, b->graph_steps);
    } else {
        // longest runs first, packed greedily into groups whose footprint
        // stays within group_bytes (a larger shot runs alone); each group
        // runs to its end before the next starts, so its fields stay in
        // the device's last-level cache between steps
        std::vector<size_t> ord(nshots);
        for (size_t k = 0; k < nshots; ++k) ord[k] = k;
        std::stable_sort(ord.begin(), ord.end(),
                         [&](size_t x, size_t y) { return nsteps[x] > nsteps[y]; });
        std::vector<E2dShot> g;
        size_t acc = 0;
        for (size_t q = 0; q <= nshots && rc == SM_OK; ++q) {
            const size_t fb = q < nshots ? shot_footprint(shots[ord[q]]) : 0;
            if (!g.empty() && (q == nshots || acc + fb > b->group_bytes)) {
                rc = b->runner.run(g, b->ctx->stream, b->graph_steps);
                g.clear();
                acc = 0;
            }
            if (q < nshots) {
                g.push_back(d[ord[q]]);
                acc += fb;
            }
        }
    }
    if (rc != SM_OK) return rc;
    SM_HIP(hipStreamSynchronize(b->ctx->stream));
    for (size_t k = 0; k < nshots; ++k) shots[k]->step += nsteps[k];
    return SM_OK;
}

int sm_e2d_batch_run_noalloc(sm_e2d_batch *b, size_t nshots, sm_e2d *const *shots,
                             const size_t *nsteps)
{
    std::vector<E2dShot> d;
    int rc = batch_shots(b, nshots, shots, nsteps, d);
    if (rc != SM_OK) return rc;
    rc = b->runner.run_static(d, b->ctx->device, b->ctx->stream);
    if (rc != SM_OK) return rc;
    SM_HIP(hipStreamSynchronize(b->ctx->stream));
    for (size_t k = 0; k < nshots; ++k) shots[k]->step += nsteps[k];
    return SM_OK;
}

int sm_e2d_batch_set_variant(sm_e2d_batch *b, int variant)
{
    SM_CHECK(b, SM_ERR_ARG, "sm_e2d_batch_set_variant: NULL batch");
    SM_CHECK(variant == SM_E2D_ROWS || variant == SM_E2D_STAGED, SM_ERR_ARG,
             "sm_e2d_batch_set_variant: unknown variant %d", variant);
    b->variant = variant;
    return SM_OK;
}

int sm_e2d_batch_set_timing(sm_e2d_batch *b, int on)
{
    SM_CHECK(b, SM_ERR_ARG, "sm_e2d_batch_set_timing: NULL batch");
    b->runner.timing = on != 0;
    b->runner.ms_total = 0.0;
    b->runner.launches = 0;
    return SM_OK;
}

int sm_e2d_batch_get_timing(const sm_e2d_batch *b, double *ms_total, size_t *launches)
{
    SM_CHECK(b && ms_total && launches, SM_ERR_ARG, "sm_e2d_batch_get_timing: NULL argument");
    *ms_total = b->runner.ms_total;
    *launches = b->runner.launches;
    return SM_OK;
}

int sm_e2d_batch_get_graphs(const sm_e2d_batch *b, size_t *live, size_t *built)
{
    SM_CHECK(b && live && built, SM_ERR_ARG, "sm_e2d_batch_get_graphs: NULL argument");
    *live = b->runner.graphs_live();
    *built = b->runner.graphs_built;
    return SM_OK;
}

}  // extern "C"
