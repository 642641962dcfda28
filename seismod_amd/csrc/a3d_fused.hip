// a3d_fused.hip -- single-pass 3D acoustic CPML step, FP32 arithmetic,
// FP32 / BF16 / FP16 storage.
//
// One launch performs one whole time step of a_iso_3d.c:599-693: the three
// velocity sweeps, the pressure sweep and the source injection, and records
// the receivers of the previous step.  The state is ping-ponged (read "in",
// write "out") so a block can recompute the few neighbour velocities it
// needs without racing the block that owns them.
//
// Work decomposition
//   - a work item is a 64 x TJ column of cells (64 lanes along z, the
//     contiguous axis; TJ waves along y) marched along x (the slowest axis)
//     over a chunk of planes [i0, i1).  A chunk with i0 > 0 first recomputes
//     vx of plane i0-1, which the pressure update of plane i0 needs.
//   - the grid is persistent: block b takes items b, b+G, b+2G, ...  Items
//     are ordered so that each XCD (blocks are dispatched round-robin over
//     the 8 XCDs) works on a contiguous range of tiles and shares halo rows
//     in its own L2.
//   - every per-plane stream (old velocities, K*dt, P and rho of the next
//     plane, neighbour rows, halos, memory variables) is loaded one plane
//     ahead into one of two alternating register sets, so HBM latency hides
//     behind the arithmetic of the current plane.  P and rho are read once
//     per plane per lane; k+1 neighbours come from the next lane.
//   - new vy of the row below comes through LDS (one barrier per plane);
//     wave 0 recomputes row j0-1, lane 0 recomputes vz at k0-1.
//   - blocks past the stencil grid copy the previous step's receiver values
//     from the "in" state, which is complete before the launch.
#include <cstdint>

#include "a3d_common.hpp"
#include "a3d_fused.hpp"
#include "seismod_hip.h"

namespace sm {

template <typename S>
struct Store;

template <>
struct Store<float> {
    __device__ static inline float ld(const float *p, size_t i, float) { return p[i]; }
    __device__ static inline void st(float *p, size_t i, float v, float) { p[i] = v; }
};

template <>
struct Store<__hip_bfloat16> {
    __device__ static inline float ld(const __hip_bfloat16 *p, size_t i, float scale)
    {
        return __bfloat162float(p[i]) * scale;
    }
    __device__ static inline void st(__hip_bfloat16 *p, size_t i, float v, float inv)
    {
        p[i] = __float2bfloat16(v * inv);
    }
};

template <>
struct Store<__half> {
    __device__ static inline float ld(const __half *p, size_t i, float scale)
    {
        return __half2float(p[i]) * scale;
    }
    __device__ static inline void st(__half *p, size_t i, float v, float inv)
    {
        p[i] = __float2half(v * inv);
    }
};

__device__ inline float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// one velocity component at one cell: v - 2dt/(rho0+rho1) * dP, with the
// CPML memory variable m (updated in place) where slab is true
__device__ inline float vel_update(float v_old, float p0, float p1, float r0, float r1, float ih,
                                   float two_dt, bool slab, float bh, float ah, float ikh, float &m)
{
    float dP = (p1 - p0) * ih;
    if (slab) {
        m = bh * m + ah * dP;
        dP = dP * ikh + m;
    }
    return v_old - two_dt * fast_rcp(r0 + r1) * dP;
}

// everything one lane reads for plane i (loaded one plane ahead)
struct Col {
    float vx, vy, vz, K;       // old velocities and K*dt at (i, j, k)
    float Pn, rn;              // P and rho at (i+1, j, k)
    float Pj, rj;              // P and rho at (i, j+1, k)
    float Pe, re;              // lane 63: (i, j, k+1) across the tile edge
    float Pjm, rjm, vyjm;      // wave 0: (i, j-1, k)
    float Pkm, rkm, vzkm;      // lane 0: (i, j, k-1)
    float mPx, mPy, mPz;       // velocity memory variables at (i, j, k)
    float mPym, mPzm;          // ... at the halo cells
    float mvx, mvy, mvz;       // pressure memory variables at (i, j, k)
};

template <typename S>
struct Ctx {
    // restrict-qualified copies of the argument pointers
    const S *__restrict__ P_in, *__restrict__ vx_in, *__restrict__ vy_in, *__restrict__ vz_in;
    S *__restrict__ P_out, *__restrict__ vx_out, *__restrict__ vy_out, *__restrict__ vz_out;
    const S *__restrict__ mPx_in, *__restrict__ mPy_in, *__restrict__ mPz_in;
    S *__restrict__ mPx_out, *__restrict__ mPy_out, *__restrict__ mPz_out;
    S *__restrict__ mvx, *__restrict__ mvy, *__restrict__ mvz;
    const float *__restrict__ rho, *__restrict__ kdt;
    int nx2, ny2, nz2, np, np2;
    size_t plane;
    float sP, isP, sV, isV, sM, isM;
    // per lane
    int lane, w, j, k, jc, kc;
    bool active;
    size_t jk;
    int syv, szv, syp, szp, syh, szh;   // slab coordinates (y/z, vel/pre, halos)
};

template <typename S>
__device__ inline void load_col(const Ctx<S> &x, int i, Col &c)
{
    const size_t o = (size_t)i * x.plane + x.jk;
    const float sP = x.sP, sV = x.sV, sM = x.sM;
    c = Col{0.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, 1.f, 0.f, 1.f, 0.f, 1.f, 0.f,
            0.f, 1.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (!x.active) return;
    c.vx = Store<S>::ld(x.vx_in, o, sV);
    c.vy = Store<S>::ld(x.vy_in, o, sV);
    c.vz = Store<S>::ld(x.vz_in, o, sV);
    c.K = x.kdt[o];
    if (i + 1 < x.nx2) {
        c.Pn = Store<S>::ld(x.P_in, o + x.plane, sP);
        c.rn = x.rho[o + x.plane];
    }
    if (x.j < x.ny2 - 1) {
        c.Pj = Store<S>::ld(x.P_in, o + x.nz2, sP);
        c.rj = x.rho[o + x.nz2];
    }
    if (x.lane == 63 && x.k + 1 < x.nz2) {
        c.Pe = Store<S>::ld(x.P_in, o + 1, sP);
        c.re = x.rho[o + 1];
    }
    if (x.w == 0 && x.j >= 1) {
        c.Pjm = Store<S>::ld(x.P_in, o - x.nz2, sP);
        c.rjm = x.rho[o - x.nz2];
        c.vyjm = Store<S>::ld(x.vy_in, o - x.nz2, sV);
        if (x.syh >= 0)
            c.mPym = Store<S>::ld(x.mPy_in, ((size_t)i * x.np2 + x.syh) * x.nz2 + x.kc, sM);
    }
    if (x.lane == 0 && x.k >= 1) {
        c.Pkm = Store<S>::ld(x.P_in, o - 1, sP);
        c.rkm = x.rho[o - 1];
        c.vzkm = Store<S>::ld(x.vz_in, o - 1, sV);
        if (x.szh >= 0)
            c.mPzm = Store<S>::ld(x.mPz_in, ((size_t)i * x.ny2 + x.jc) * x.np2 + x.szh, sM);
    }
    const int sxv = slab_vel(i, x.nx2, x.np);
    if (sxv >= 0 && i < x.nx2 - 1)
        c.mPx = Store<S>::ld(x.mPx_in, ((size_t)sxv * x.ny2 + x.jc) * x.nz2 + x.kc, sM);
    if (x.syv >= 0) c.mPy = Store<S>::ld(x.mPy_in, ((size_t)i * x.np2 + x.syv) * x.nz2 + x.kc, sM);
    if (x.szv >= 0) c.mPz = Store<S>::ld(x.mPz_in, ((size_t)i * x.ny2 + x.jc) * x.np2 + x.szv, sM);
    const bool pin = i >= 1 && x.j >= 1 && x.k >= 1;
    const int sxp = slab_pre(i, x.nx2, x.np);
    if (pin && sxp >= 0) c.mvx = Store<S>::ld(x.mvx, ((size_t)sxp * x.ny2 + x.jc) * x.nz2 + x.kc, sM);
    if (pin && x.syp >= 0) c.mvy = Store<S>::ld(x.mvy, ((size_t)i * x.np2 + x.syp) * x.nz2 + x.kc, sM);
    if (pin && x.szp >= 0) c.mvz = Store<S>::ld(x.mvz, ((size_t)i * x.ny2 + x.jc) * x.np2 + x.szp, sM);
}

// one plane: load the next plane's column into nxt, update plane i from cur
template <typename S, int TJ>
__device__ __forceinline__ void plane_step(const Ctx<S> &x, const FusedArgs<S> &a, int i, int i0,
                                           int i1, Col &cur, Col &nxt, float &P_c, float &r_c,
                                           float &vx_prev, float (*s_vy)[TJ][64])
{
    if (i + 1 < i1) load_col<S>(x, i + 1, nxt);
    const bool own = i >= i0;
    const size_t o = (size_t)i * x.plane + x.jk;
    const float isM = x.isM, isV = x.isV;

    // vx at (i, j, k): a_iso_3d.c:605-619
    float vxn = cur.vx;
    if (x.active && i < x.nx2 - 1) {
        const int sxv = slab_vel(i, x.nx2, x.np);
        float m = cur.mPx;
        vxn = vel_update(cur.vx, P_c, cur.Pn, r_c, cur.rn, a.idx, a.two_dt, sxv >= 0, a.bhx[i],
                         a.ahx[i], a.ikhx[i], m);
        if (own) {
            Store<S>::st(x.vx_out, o, vxn, isV);
            if (sxv >= 0)
                Store<S>::st(x.mPx_out, ((size_t)sxv * x.ny2 + x.jc) * x.nz2 + x.kc, m, isM);
        }
    }
    if (!own) {   // chunk start: only vx of the plane below is needed
        vx_prev = vxn;
        P_c = cur.Pn;
        r_c = cur.rn;
        return;
    }

    // vy, vz at (i, j, k): a_iso_3d.c:621-651
    float P_kp = __shfl_down(P_c, 1);
    float r_kp = __shfl_down(r_c, 1);
    if (x.lane == 63) { P_kp = cur.Pe; r_kp = cur.re; }
    float vyn = cur.vy, vzn = cur.vz;
    if (x.active && x.j < x.ny2 - 1) {
        float m = cur.mPy;
        const int n = x.j;
        vyn = vel_update(cur.vy, P_c, cur.Pj, r_c, cur.rj, a.idy, a.two_dt, x.syv >= 0, a.bhy[n],
                         a.ahy[n], a.ikhy[n], m);
        Store<S>::st(x.vy_out, o, vyn, isV);
        if (x.syv >= 0) Store<S>::st(x.mPy_out, ((size_t)i * x.np2 + x.syv) * x.nz2 + x.kc, m, isM);
    }
    if (x.active && x.k < x.nz2 - 1) {
        float m = cur.mPz;
        const int n = x.k;
        vzn = vel_update(cur.vz, P_c, P_kp, r_c, r_kp, a.idz, a.two_dt, x.szv >= 0, a.bhz[n],
                         a.ahz[n], a.ikhz[n], m);
        Store<S>::st(x.vz_out, o, vzn, isV);
        if (x.szv >= 0) Store<S>::st(x.mPz_out, ((size_t)i * x.ny2 + x.jc) * x.np2 + x.szv, m, isM);
    }
    // halo velocities owned by neighbouring tiles (recomputed, not stored)
    float vy_halo = 0.f, vz_halo = 0.f;
    if (x.w == 0 && x.active && x.j >= 1) {
        float m = cur.mPym;
        const int n = x.j - 1;
        vy_halo = vel_update(cur.vyjm, cur.Pjm, P_c, cur.rjm, r_c, a.idy, a.two_dt, x.syh >= 0,
                             a.bhy[n], a.ahy[n], a.ikhy[n], m);
    }
    if (x.lane == 0 && x.active && x.k >= 1) {
        float m = cur.mPzm;
        const int n = x.k - 1;
        vz_halo = vel_update(cur.vzkm, cur.Pkm, P_c, cur.rkm, r_c, a.idz, a.two_dt, x.szh >= 0,
                             a.bhz[n], a.ahz[n], a.ikhz[n], m);
    }
    s_vy[i & 1][x.w][x.lane] = vyn;
    __syncthreads();
    const float vy_m = x.w > 0 ? s_vy[i & 1][x.w - 1][x.lane] : vy_halo;
    float vz_m = __shfl_up(vzn, 1);
    if (x.lane == 0) vz_m = vz_halo;

    // pressure at (i, j, k): a_iso_3d.c:659-683
    if (x.active) {
        float Pn = P_c;
        if (i >= 1 && x.j >= 1 && x.k >= 1) {
            float dvdx = (vxn - vx_prev) * a.idx;
            float dvdy = (vyn - vy_m) * a.idy;
            float dvdz = (vzn - vz_m) * a.idz;
            const int sxp = slab_pre(i, x.nx2, x.np);
            if (sxp >= 0) {
                const float m = a.bx[i] * cur.mvx + a.ax[i] * dvdx;
                Store<S>::st(x.mvx, ((size_t)sxp * x.ny2 + x.jc) * x.nz2 + x.kc, m, isM);
                dvdx = dvdx * a.ikx[i] + m;
            }
            if (x.syp >= 0) {
                const float m = a.by[x.j] * cur.mvy + a.ay[x.j] * dvdy;
                Store<S>::st(x.mvy, ((size_t)i * x.np2 + x.syp) * x.nz2 + x.kc, m, isM);
                dvdy = dvdy * a.iky[x.j] + m;
            }
            if (x.szp >= 0) {
                const float m = a.bz[x.k] * cur.mvz + a.az[x.k] * dvdz;
                Store<S>::st(x.mvz, ((size_t)i * x.ny2 + x.jc) * x.np2 + x.szp, m, isM);
                dvdz = dvdz * a.ikz[x.k] + m;
            }
            Pn = P_c - cur.K * (dvdx + dvdy + dvdz);
        }
        // sources of this plane (a_iso_3d.c:689-693), in source order
        const int q0 = a.src_plane[i], q1 = a.src_plane[i + 1];
        for (int q = q0; q < q1; ++q)
            if ((size_t)a.src_cell[q] == o && a.step < a.src_len[q])
                Pn += a.src_inc[a.src_off[q] + a.step];
        Store<S>::st(x.P_out, o, Pn, x.isP);
    }
    vx_prev = vxn;
    P_c = cur.Pn;
    r_c = cur.rn;
}

template <typename S, int TJ>
__global__ void __launch_bounds__(64 * TJ)
a3d_fused_step(FusedArgs<S> a)
{
    __shared__ float s_vy[2][TJ][64];

    // ---------------- receiver gather blocks (previous step) -------------
    if ((int)blockIdx.x >= a.nblocks) {
        if (a.gather_row < 0) return;
        const int r = ((int)blockIdx.x - a.nblocks) * (64 * TJ) + threadIdx.y * 64 + threadIdx.x;
        if (r < a.nrec) {
            const int cp = a.rcomp[r];
            const S *f = cp == SM_VX ? a.vx_in : cp == SM_VY ? a.vy_in : cp == SM_VZ ? a.vz_in : a.P_in;
            const float sc = cp == SM_P ? a.sP : a.sV;
            a.trace[(size_t)a.gather_row * a.nrec + r] = Store<S>::ld(f, (size_t)a.ridx[r], sc);
        }
        return;
    }

    Ctx<S> x;
    x.P_in = a.P_in; x.vx_in = a.vx_in; x.vy_in = a.vy_in; x.vz_in = a.vz_in;
    x.P_out = a.P_out; x.vx_out = a.vx_out; x.vy_out = a.vy_out; x.vz_out = a.vz_out;
    x.mPx_in = a.mPx_in; x.mPy_in = a.mPy_in; x.mPz_in = a.mPz_in;
    x.mPx_out = a.mPx_out; x.mPy_out = a.mPy_out; x.mPz_out = a.mPz_out;
    x.mvx = a.mvx; x.mvy = a.mvy; x.mvz = a.mvz;
    x.rho = a.rho; x.kdt = a.kdt;
    x.nx2 = a.nx2; x.ny2 = a.ny2; x.nz2 = a.nz2; x.np = a.np; x.np2 = 2 * a.np;
    x.plane = (size_t)a.ny2 * a.nz2;
    x.sP = a.sP; x.isP = a.isP; x.sV = a.sV; x.isV = a.isV; x.sM = a.sM; x.isM = a.isM;
    x.lane = threadIdx.x;
    x.w = threadIdx.y;

    for (int item = blockIdx.x; item < a.nitems; item += a.nblocks) {
        // XCD-contiguous order: items of one XCD (item % 8) are consecutive
        // logical tiles.  nitems is a multiple of 8; the padding items idle.
        const int logical = (item % 8) * (a.nitems / 8) + item / 8;
        __syncthreads();   // LDS of the previous item is no longer read
        if (logical >= a.nreal) continue;   // uniform per block
        const int chunk = logical / a.ntiles;
        const int tile = logical - chunk * a.ntiles;
        const int tj = tile / a.ntk;
        const int tk = tile - tj * a.ntk;
        x.k = tk * 64 + x.lane;
        x.j = tj * TJ + x.w;
        x.active = x.k < x.nz2 && x.j < x.ny2;
        x.kc = x.active ? x.k : 0;
        x.jc = x.active ? x.j : 0;
        x.jk = (size_t)x.jc * x.nz2 + x.kc;
        x.syv = x.active && x.j < x.ny2 - 1 ? slab_vel(x.j, x.ny2, x.np) : -1;
        x.szv = x.active && x.k < x.nz2 - 1 ? slab_vel(x.k, x.nz2, x.np) : -1;
        x.syp = x.active ? slab_pre(x.j, x.ny2, x.np) : -1;
        x.szp = x.active ? slab_pre(x.k, x.nz2, x.np) : -1;
        x.syh = x.active && x.j >= 1 ? slab_vel(x.j - 1, x.ny2, x.np) : -1;
        x.szh = x.active && x.k >= 1 ? slab_vel(x.k - 1, x.nz2, x.np) : -1;
        const int i0 = chunk * a.chunk_len;
        const int i1 = min(x.nx2, i0 + a.chunk_len);
        if (i0 >= x.nx2) continue;   // uniform; the host never makes empty chunks
        const int ist = i0 > 0 ? i0 - 1 : 0;

        float P_c = 0.f, r_c = 1.f, vx_prev = 0.f;
        if (x.active) {
            const size_t o = (size_t)ist * x.plane + x.jk;
            P_c = Store<S>::ld(x.P_in, o, x.sP);
            r_c = x.rho[o];
        }
        Col A, B;
        load_col<S>(x, ist, A);
        for (int i = ist; i < i1; i += 2) {
            plane_step<S, TJ>(x, a, i, i0, i1, A, B, P_c, r_c, vx_prev, s_vy);
            if (i + 1 >= i1) break;
            plane_step<S, TJ>(x, a, i + 1, i0, i1, B, A, P_c, r_c, vx_prev, s_vy);
        }
    }
}

template <typename S>
__global__ void __launch_bounds__(256)
a3d_gather(FusedArgs<S> a)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (a.gather_row < 0 || r >= a.nrec) return;
    const int cp = a.rcomp[r];
    const S *f = cp == SM_VX ? a.vx_in : cp == SM_VY ? a.vy_in : cp == SM_VZ ? a.vz_in : a.P_in;
    const float sc = cp == SM_P ? a.sP : a.sV;
    a.trace[(size_t)a.gather_row * a.nrec + r] = Store<S>::ld(f, (size_t)a.ridx[r], sc);
}

template <typename S, int TJ>
static const void *kernel_ptr()
{
    return reinterpret_cast<const void *>(&a3d_fused_step<S, TJ>);
}

template <typename S>
int fused_max_blocks_per_cu(int tj, int *blocks)
{
    const void *f = nullptr;
    switch (tj) {
    case 1: f = kernel_ptr<S, 1>(); break;
    case 2: f = kernel_ptr<S, 2>(); break;
    case 4: f = kernel_ptr<S, 4>(); break;
    case 8: f = kernel_ptr<S, 8>(); break;
    default: return SM_ERR_ARG;
    }
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64 * tj, 0) != hipSuccess)
        return SM_ERR_HIP;
    *blocks = nb;
    return SM_OK;
}

template <typename S>
int fused_launch(const FusedArgs<S> &a, int tj, hipStream_t st)
{
    const int blocks = a.nblocks + a.gather_blocks;
    switch (tj) {
    case 1: a3d_fused_step<S, 1><<<blocks, dim3(64, 1), 0, st>>>(a); break;
    case 2: a3d_fused_step<S, 2><<<blocks, dim3(64, 2), 0, st>>>(a); break;
    case 4: a3d_fused_step<S, 4><<<blocks, dim3(64, 4), 0, st>>>(a); break;
    case 8: a3d_fused_step<S, 8><<<blocks, dim3(64, 8), 0, st>>>(a); break;
    default: return SM_ERR_ARG;
    }
    return hipGetLastError() == hipSuccess ? SM_OK : SM_ERR_HIP;
}

template <typename S>
int gather_launch(const FusedArgs<S> &a, hipStream_t st)
{
    if (a.nrec == 0 || a.gather_row < 0) return SM_OK;
    a3d_gather<S><<<(a.nrec + 255) / 256, 256, 0, st>>>(a);
    return hipGetLastError() == hipSuccess ? SM_OK : SM_ERR_HIP;
}

template int fused_launch<float>(const FusedArgs<float> &, int, hipStream_t);
template int fused_launch<__hip_bfloat16>(const FusedArgs<__hip_bfloat16> &, int, hipStream_t);
template int fused_launch<__half>(const FusedArgs<__half> &, int, hipStream_t);
template int gather_launch<float>(const FusedArgs<float> &, hipStream_t);
template int gather_launch<__hip_bfloat16>(const FusedArgs<__hip_bfloat16> &, hipStream_t);
template int gather_launch<__half>(const FusedArgs<__half> &, hipStream_t);
template int fused_max_blocks_per_cu<float>(int, int *);
template int fused_max_blocks_per_cu<__hip_bfloat16>(int, int *);
template int fused_max_blocks_per_cu<__half>(int, int *);

}  // namespace sm
