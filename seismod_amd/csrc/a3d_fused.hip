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
#include "a3d_store.hpp"
#include "seismod_hip.h"

namespace sm {

// what one lane reads for plane i, loaded one plane ahead
struct Col {
    float vx, vy, vz, K;    // old velocities and K*dt at (i, j, k)
    float Pn, rn;           // P, rho at (i+1, j, k)   (clamped to the last plane)
    float Pj, rj;           // P, rho at (i, j+1, k)   (clamped to the last row)
    float Pk, rk;           // P, rho at (i, j, k+1)   (clamped to the last column)
    float Pjm, rjm, vyjm;   // wave 0 only: (i, j-1, k)
    float mPx, mPy, mPz;    // velocity memory variables (slab cells only)
    float mPym;             // wave 0 only: at the halo row
    float mvx, mvy, mvz;    // pressure memory variables (slab cells only)
};

// constants of one lane for one work item
struct Lane {
    unsigned o_jk, o_jp, o_kp, o_jm;   // in-plane offsets of the cell and its neighbours
    int j, kk, kc;                     // j is wave-uniform
    bool own;                          // this lane writes its cell
    int syv, syp, syh;                 // y slab coordinates (wave-uniform)
    int szv, szp;                      // z slab coordinates (per lane)
    unsigned mz_v, mz_p;               // z-slab offsets within a plane
    // CPML coefficients along y (wave-uniform) and z (per lane)
    float ybh, yah, yikh, yb, ya, yik, hbh, hah, hikh;
    float zbh, zah, zikh, zb, za, zik;
};

template <typename S>
__device__ __forceinline__ void load_col(const FusedArgs<S> &a, const Lane &L, int w, int i,
                                         unsigned plane, int sxv, int sxp, Col &c)
{
    const unsigned o = (unsigned)i * plane;
    const unsigned on = (unsigned)min(i + 1, a.nx2 - 1) * plane;
    const unsigned np2 = 2u * a.np;
    c.vx = Store<S>::ld(a.in, F_VX * a.nnn + o + L.o_jk, a.sV);
    c.vy = Store<S>::ld(a.in, F_VY * a.nnn + o + L.o_jk, a.sV);
    c.vz = Store<S>::ld(a.in, F_VZ * a.nnn + o + L.o_jk, a.sV);
    c.K = ld_at(a.mat, a.nnn + o + L.o_jk);
    c.Pn = Store<S>::ld(a.in, on + L.o_jk, a.sP);
    c.rn = ld_at(a.mat, on + L.o_jk);
    c.Pj = Store<S>::ld(a.in, o + L.o_jp, a.sP);
    c.rj = ld_at(a.mat, o + L.o_jp);
    c.Pk = Store<S>::ld(a.in, o + L.o_kp, a.sP);
    c.rk = ld_at(a.mat, o + L.o_kp);
    if (w == 0) {
        c.Pjm = Store<S>::ld(a.in, o + L.o_jm, a.sP);
        c.rjm = ld_at(a.mat, o + L.o_jm);
        c.vyjm = Store<S>::ld(a.in, F_VY * a.nnn + o + L.o_jm, a.sV);
        if (L.syh >= 0)
            c.mPym = Store<S>::ld(a.mem_in, a.my + ((unsigned)i * np2 + L.syh) * a.pz + L.kc, a.sMv);
    }
    if (sxv >= 0) c.mPx = Store<S>::ld(a.mem_in, (unsigned)sxv * plane + L.o_jk, a.sMv);
    if (L.syv >= 0)
        c.mPy = Store<S>::ld(a.mem_in, a.my + ((unsigned)i * np2 + L.syv) * a.pz + L.kc, a.sMv);
    if (L.szv >= 0) c.mPz = Store<S>::ld(a.mem_in, a.mz + (unsigned)i * a.ny2 * np2 + L.mz_v, a.sMv);
    if (sxp >= 0) c.mvx = Store<S>::ld(a.memp, (unsigned)sxp * plane + L.o_jk, a.sMp);
    if (L.syp >= 0)
        c.mvy = Store<S>::ld(a.memp, a.my + ((unsigned)i * np2 + L.syp) * a.pz + L.kc, a.sMp);
    if (L.szp >= 0) c.mvz = Store<S>::ld(a.memp, a.mz + (unsigned)i * a.ny2 * np2 + L.mz_p, a.sMp);
}

// one plane: prefetch plane i+PF into fill, update plane i from cur
template <typename S, int TJ, int PF>
__device__ __forceinline__ void plane_step(const FusedArgs<S> &a, const Lane &L, int w, int i,
                                           int i0, int i1, unsigned plane, Col &cur, Col &fill,
                                           float &P_c, float &r_c, float &vx_prev,
                                           float (*s_vy)[TJ][64])
{
    if (i + PF < i1) {
        const XPlane xn = xplane(a.xpl, i + PF);
        load_col<S>(a, L, w, i + PF, plane, xn.sxv, xn.sxp, fill);
    }
    const XPlane x = xplane(a.xpl, i);
    const unsigned o = (unsigned)i * plane + L.o_jk;
    const unsigned np2 = 2u * a.np;

    // vx at (i, j, k), a_iso_3d.c:605-619.  On the last plane the clamped
    // neighbour equals P_c, so dP = 0 and vx keeps its value.
    float mx = cur.mPx;
    const float vxn = vel_update(cur.vx, P_c, cur.Pn, r_c, cur.rn, a.idx, a.two_dt, x.sxv >= 0,
                                 x.bh, x.ah, x.ikh, mx);
    if (i < i0) {   // chunk start: only vx of the plane below is needed
        vx_prev = vxn;
        P_c = cur.Pn;
        r_c = cur.rn;
        return;
    }
    // vy, vz at (i, j, k), a_iso_3d.c:621-651
    float my = cur.mPy, mz = cur.mPz;
    const float vyn = vel_update(cur.vy, P_c, cur.Pj, r_c, cur.rj, a.idy, a.two_dt, L.syv >= 0,
                                 L.ybh, L.yah, L.yikh, my);
    const float vzn = vel_update(cur.vz, P_c, cur.Pk, r_c, cur.rk, a.idz, a.two_dt, L.szv >= 0,
                                 L.zbh, L.zah, L.zikh, mz);
    if (L.own) {
        Store<S>::st(a.out, F_VX * a.nnn + o, vxn, a.isV, a.salt, a.sr);
        Store<S>::st(a.out, F_VY * a.nnn + o, vyn, a.isV, a.salt, a.sr);
        Store<S>::st(a.out, F_VZ * a.nnn + o, vzn, a.isV, a.salt, a.sr);
        if (x.sxv >= 0)
            Store<S>::st(a.mem_out, (unsigned)x.sxv * plane + L.o_jk, mx, a.isMv,
                         a.salt ^ SALT_MV, a.sr);
        if (L.syv >= 0)
            Store<S>::st(a.mem_out, a.my + ((unsigned)i * np2 + L.syv) * a.pz + L.kc, my, a.isMv,
                         a.salt ^ SALT_MV, a.sr);
        if (L.szv >= 0)
            Store<S>::st(a.mem_out, a.mz + (unsigned)i * a.ny2 * np2 + L.mz_v, mz, a.isMv,
                         a.salt ^ SALT_MV, a.sr);
    }
    // vy of the row below the tile (owned by the tile below; not stored)
    float vy_halo = 0.f;
    if (w == 0) {
        float mh = cur.mPym;
        vy_halo = vel_update(cur.vyjm, cur.Pjm, P_c, cur.rjm, r_c, a.idy, a.two_dt, L.syh >= 0,
                             L.hbh, L.hah, L.hikh, mh);
    }
    s_vy[i & 1][w][threadIdx.x] = vyn;
    __syncthreads();
    const float vy_m = w > 0 ? s_vy[i & 1][w - 1][threadIdx.x] : vy_halo;
    const float vz_m = __shfl_up(vzn, 1);   // lane 0 is the left halo lane

    // pressure at (i, j, k), a_iso_3d.c:659-683
    if (L.own) {
        float Pn = P_c;
        if (i >= 1 && L.j >= 1 && L.kk >= 1) {
            float dvdx = (vxn - vx_prev) * a.idx;
            float dvdy = (vyn - vy_m) * a.idy;
            float dvdz = (vzn - vz_m) * a.idz;
            if (x.sxp >= 0) {
                const float m = x.b * cur.mvx + x.a * dvdx;
                Store<S>::st(a.memp, (unsigned)x.sxp * plane + L.o_jk, m, a.isMp,
                             a.salt ^ SALT_MP, a.sr);
                dvdx = dvdx * x.ik + m;
            }
            if (L.syp >= 0) {
                const float m = L.yb * cur.mvy + L.ya * dvdy;
                Store<S>::st(a.memp, a.my + ((unsigned)i * np2 + L.syp) * a.pz + L.kc, m, a.isMp,
                             a.salt ^ SALT_MP, a.sr);
                dvdy = dvdy * L.yik + m;
            }
            if (L.szp >= 0) {
                const float m = L.zb * cur.mvz + L.za * dvdz;
                Store<S>::st(a.memp, a.mz + (unsigned)i * a.ny2 * np2 + L.mz_p, m, a.isMp,
                             a.salt ^ SALT_MP, a.sr);
                dvdz = dvdz * L.zik + m;
            }
            Pn = P_c - cur.K * (dvdx + dvdy + dvdz);
        }
        // sources of this plane (a_iso_3d.c:689-693), in source order
        for (int q = x.q0; q < x.q1; ++q)
            if ((unsigned)a.src[q] == o && a.step < a.src[2 * a.nsrc + q])
                Pn += a.src_inc[a.src[a.nsrc + q] + a.step];
        Store<S>::st(a.out, o, Pn, a.isP, a.salt, a.sr);
    }
    vx_prev = vxn;
    P_c = cur.Pn;
    r_c = cur.rn;
}

template <typename S>
__device__ __forceinline__ void gather_one(const FusedArgs<S> &a, int r)
{
    const int cp = a.rcomp[r];
    const unsigned q = cp == SM_VX ? 1u : cp == SM_VY ? 2u : cp == SM_VZ ? 3u : 0u;
    const float sc = cp == SM_P ? a.sP : a.sV;
    a.trace[(size_t)a.gather_row * a.nrec + r] = Store<S>::ld(a.in, q * a.nnn + (unsigned)a.ridx[r], sc);
}

// PF: planes prefetched ahead of the one being updated (1 or 2).  Tiles of
// 63 owned cells along z; lane 0 holds the left neighbour.
template <typename S, int TJ, int PF>
__global__ void __launch_bounds__(64 * TJ)
a3d_fused_step(FusedArgs<S> a)
{
    __shared__ float s_vy[2][TJ][64];

    // ---------------- receiver gather blocks (previous step) -------------
    if ((int)blockIdx.x >= a.nblocks) {
        if (a.gather_row < 0) return;
        const int r = ((int)blockIdx.x - a.nblocks) * (64 * TJ) + threadIdx.y * 64 + threadIdx.x;
        if (r < a.nrec) gather_one(a, r);
        return;
    }

    const int lane = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.y);   // one wave per y row
    const unsigned plane = (unsigned)a.ny2 * a.pz;
    const unsigned np2 = 2u * a.np;

    for (int item = blockIdx.x; item < a.nitems; item += a.nblocks) {
        // XCD-contiguous order: items of one XCD (item % 8) are consecutive
        // logical tiles.  nitems is a multiple of 8; padding items idle.
        const int logical = (item % 8) * (a.nitems / 8) + item / 8;
        __syncthreads();   // the LDS of the previous item is no longer read
        if (logical >= a.nreal) continue;   // uniform per block
        const int chunk = logical / a.ntiles;
        const int tile = logical - chunk * a.ntiles;
        const int tj = tile / a.ntk;
        const int tk = tile - tj * a.ntk;
        // 63 owned cells per wave along z; lane 0 sits on the last cell of
        // the tile to the left and only supplies vz at k0-1.  Cells past the
        // grid edge read the last cell (its neighbour differences vanish)
        // and are not owned.
        Lane L;
        const int k0 = tk * 63;
        L.kk = k0 - 1 + lane;
        L.j = tj * TJ + w;
        const bool jv = L.j < a.ny2, kv = L.kk >= 0 && L.kk < a.nz2;
        L.own = jv && kv && lane >= 1;
        const int jc = min(L.j, a.ny2 - 1);
        L.kc = min(max(L.kk, 0), a.nz2 - 1);
        const int jp = min(jc + 1, a.ny2 - 1), kp = min(L.kc + 1, a.nz2 - 1), jm = max(jc - 1, 0);
        L.o_jk = (unsigned)jc * a.pz + L.kc;
        L.o_jp = (unsigned)jp * a.pz + L.kc;
        L.o_kp = (unsigned)jc * a.pz + kp;
        L.o_jm = (unsigned)jm * a.pz + L.kc;
        L.syv = jv && L.j < a.ny2 - 1 ? slab_vel(L.j, a.ny2, a.np) : -1;
        L.syp = jv ? slab_pre(L.j, a.ny2, a.np) : -1;
        L.syh = jv && L.j >= 1 ? slab_vel(L.j - 1, a.ny2, a.np) : -1;
        L.szv = kv && L.kk < a.nz2 - 1 ? slab_vel(L.kk, a.nz2, a.np) : -1;
        L.szp = kv ? slab_pre(L.kk, a.nz2, a.np) : -1;
        L.mz_v = (unsigned)jc * np2 + (L.szv < 0 ? 0 : L.szv);
        L.mz_p = (unsigned)jc * np2 + (L.szp < 0 ? 0 : L.szp);

        {
            // y coefficients of the row (wave-uniform) and of the halo row,
            // z coefficients of the lane's column; tables a, b, 1/k, ah, bh, 1/kh
            const float *ty = a.tab + a.toff_y, *tz = a.tab + a.toff_z;
            const int ny2 = a.ny2, nz2 = a.nz2;
            const int jn = max(min(L.j, ny2 - 2), 0), jh = max(min(L.j - 1, ny2 - 2), 0);
            const int kn = max(min(L.kc, nz2 - 2), 0);
            L.ybh = ty[4 * ny2 + jn]; L.yah = ty[3 * ny2 + jn]; L.yikh = ty[5 * ny2 + jn];
            L.hbh = ty[4 * ny2 + jh]; L.hah = ty[3 * ny2 + jh]; L.hikh = ty[5 * ny2 + jh];
            L.yb = ty[ny2 + jc]; L.ya = ty[jc]; L.yik = ty[2 * ny2 + jc];
            L.zbh = ld_at(tz, 4 * nz2 + kn); L.zah = ld_at(tz, 3 * nz2 + kn);
            L.zikh = ld_at(tz, 5 * nz2 + kn);
            L.zb = ld_at(tz, nz2 + L.kc); L.za = ld_at(tz, L.kc); L.zik = ld_at(tz, 2 * nz2 + L.kc);
        }
        const int i0 = chunk * a.chunk_len;
        const int i1 = min(a.nx2, i0 + a.chunk_len);
        if (i0 >= a.nx2) continue;   // uniform; the host never makes empty chunks
        const int ist = i0 > 0 ? i0 - 1 : 0;

        float P_c = Store<S>::ld(a.in, (unsigned)ist * plane + L.o_jk, a.sP);
        float r_c = ld_at(a.mat, (unsigned)ist * plane + L.o_jk);
        float vx_prev = 0.f;
        Col A, B;
        A.Pjm = A.rjm = A.vyjm = A.mPym = 0.f;
        A.mPx = A.mPy = A.mPz = A.mvx = A.mvy = A.mvz = 0.f;
        A.Pk = A.rk = 0.f;
        B = A;
        {
            const XPlane xs = xplane(a.xpl, ist);
            load_col<S>(a, L, w, ist, plane, xs.sxv, xs.sxp, A);
        }
        if constexpr (PF == 1) {
            for (int i = ist; i < i1; i += 2) {
                plane_step<S, TJ, 1>(a, L, w, i, i0, i1, plane, A, B, P_c, r_c, vx_prev, s_vy);
                if (i + 1 >= i1) break;
                plane_step<S, TJ, 1>(a, L, w, i + 1, i0, i1, plane, B, A, P_c, r_c, vx_prev, s_vy);
            }
        } else {
            // three register sets in rotation: plane i in use, i+1 and i+2 in flight
            Col C = B;   // B still holds the zeroed template (its loads follow)
            if (ist + 1 < i1) {
                const XPlane xs = xplane(a.xpl, ist + 1);
                load_col<S>(a, L, w, ist + 1, plane, xs.sxv, xs.sxp, B);
            }
            for (int i = ist; i < i1; i += 3) {
                plane_step<S, TJ, 2>(a, L, w, i, i0, i1, plane, A, C, P_c, r_c, vx_prev, s_vy);
                if (i + 1 >= i1) break;
                plane_step<S, TJ, 2>(a, L, w, i + 1, i0, i1, plane, B, A, P_c, r_c, vx_prev, s_vy);
                if (i + 2 >= i1) break;
                plane_step<S, TJ, 2>(a, L, w, i + 2, i0, i1, plane, C, B, P_c, r_c, vx_prev, s_vy);
            }
        }
    }
}

template <typename S>
__global__ void __launch_bounds__(256)
a3d_gather(FusedArgs<S> a)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (a.gather_row < 0 || r >= a.nrec) return;
    gather_one(a, r);
}

template <typename S, int TJ, int PF>
static const void *kernel_ptr()
{
    return reinterpret_cast<const void *>(&a3d_fused_step<S, TJ, PF>);
}

// the instantiated (TJ, PF) combinations, as one switch key
#define SM_FUSED_CASES(X) X(1, 1) X(2, 1) X(4, 1) X(8, 1) X(1, 2) X(2, 2) X(4, 2) X(8, 2)
#define SM_FUSED_KEY(tj, pf) ((tj) * 10 + (pf))

template <typename S>
int fused_max_blocks_per_cu(int tj, int pf, int *blocks)
{
    const void *f = nullptr;
    switch (SM_FUSED_KEY(tj, pf)) {
#define SM_X(T, P) case SM_FUSED_KEY(T, P): f = kernel_ptr<S, T, P>(); break;
        SM_FUSED_CASES(SM_X)
#undef SM_X
    default: return SM_ERR_ARG;
    }
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64 * tj, 0) != hipSuccess)
        return SM_ERR_HIP;
    *blocks = nb;
    return SM_OK;
}

template <typename S>
int fused_launch(const FusedArgs<S> &a, int tj, int pf, hipStream_t st)
{
    const int blocks = a.nblocks + a.gather_blocks;
    switch (SM_FUSED_KEY(tj, pf)) {
#define SM_X(T, P)                                                                     \
    case SM_FUSED_KEY(T, P):                                                           \
        a3d_fused_step<S, T, P><<<blocks, dim3(64, T), 0, st>>>(a);                    \
        break;
        SM_FUSED_CASES(SM_X)
#undef SM_X
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

template int fused_launch<float>(const FusedArgs<float> &, int, int, hipStream_t);
template int fused_launch<__hip_bfloat16>(const FusedArgs<__hip_bfloat16> &, int, int, hipStream_t);
template int fused_launch<__half>(const FusedArgs<__half> &, int, int, hipStream_t);
template int gather_launch<float>(const FusedArgs<float> &, hipStream_t);
template int gather_launch<__hip_bfloat16>(const FusedArgs<__hip_bfloat16> &, hipStream_t);
template int gather_launch<__half>(const FusedArgs<__half> &, hipStream_t);
template int fused_max_blocks_per_cu<float>(int, int, int *);
template int fused_max_blocks_per_cu<__hip_bfloat16>(int, int, int *);
template int fused_max_blocks_per_cu<__half>(int, int, int *);

}  // namespace sm
