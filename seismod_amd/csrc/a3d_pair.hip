// a3d_pair.hip -- single-pass 3D acoustic CPML step with two z-adjacent
// cells per lane (FP32 arithmetic; FP32 / BF16 / FP16 storage).
//
// Same time step and the same per-cell arithmetic as a3d_fused_step (so
// the results are bit-identical to it), with a memory-access shape that
// moves twice the bytes per instruction:
//   - a wave covers two y rows (lanes 0-31 row j, lanes 32-63 row j+1) and
//     64 cells of z per row, each lane holding cells k0 = kb + 2m and k0+1
//     (m = lane % 32) loaded and stored as one pair: dwordx2 for FP32, one
//     dword of two 16-bit values for BF16/FP16;
//   - a z tile starts at an even cell kb = 62 t: the first cell (kb) is
//     the left halo (it supplies vz at kb for the pressure of kb+1), cells
//     kb+1 .. kb+62 are owned, the last (kb+63) is the right halo (it
//     supplies P and rho at kb+63).  Tile 0 also owns cell 0, whose pressure
//     never changes (the pressure sweep starts at k = 1, a_iso_3d.c:660).
//     The padded z extent must be even so that pairs never straddle rows;
//   - z neighbours move between the two cells of a lane or to the next lane
//     by shuffles; the row above the lower half is the upper half (shuffle
//     by 32 lanes); the upper half loads its row above; new vy of the row
//     below the lower half comes from the wave below through LDS (one
//     barrier per plane), or for wave 0 is recomputed from that row;
//   - work items, x-march, one-plane-ahead prefetch, persistent grid,
//     XCD-contiguous item order and the receiver gather blocks are those of
//     a3d_fused_step.
#include <cstdint>

#include "a3d_common.hpp"
#include "a3d_fused.hpp"
#include "a3d_store.hpp"
#include "seismod_hip.h"

namespace sm {
namespace {

// what one lane reads for plane i (two cells), loaded one plane ahead and
// kept as stored (packed for 16-bit storage) until used
template <typename S>
struct PCol {
    using R = typename Pair<S>::Raw;
    R vx, vy, vz;           // old velocities at (i, j, k0..k0+1)
    R Pn;                   // P at (i+1, ...)   (clamped to the last plane)
    R Pj;                   // upper half: P at (i, j+1, ...) (clamped to the last row)
    R Pjm, vyjm;            // wave 0, lower half: P, vy at (i, j-1, ...)
    float2 K, rn, rj, rjm;  // K*dt at (i, ...); rho at (i+1, ...), (i, j+1, ...), (i, j-1, ...)
};

struct PLane {
    int j;                  // row of this half (may exceed ny2-1: then not owned)
    int h;                  // 0 lower half, 1 upper half
    int k0;                 // first cell of the lane's pair
    bool own0, own1;        // the lane writes cell k0 / k0+1
    unsigned o, o_jp, o_jm; // in-plane offsets of the pair, its row above and below (even)
};

template <typename S, int TJ>
__device__ __forceinline__ void pload(const FusedArgs<S> &a, const PLane &L, int w, int i,
                                      unsigned plane, PCol<S> &c)
{
    const unsigned o = (unsigned)i * plane;
    const unsigned on = (unsigned)min(i + 1, a.nx2 - 1) * plane;
    c.vx = Pair<S>::raw(a.in, F_VX * a.nnn + o + L.o);
    c.vy = Pair<S>::raw(a.in, F_VY * a.nnn + o + L.o);
    c.vz = Pair<S>::raw(a.in, F_VZ * a.nnn + o + L.o);
    c.K = Pair<float>::ld(a.mat, a.nnn + o + L.o, 1.f);
    c.Pn = Pair<S>::raw(a.in, on + L.o);
    c.rn = Pair<float>::ld(a.mat, on + L.o, 1.f);
    if (L.h) {
        c.Pj = Pair<S>::raw(a.in, o + L.o_jp);
        c.rj = Pair<float>::ld(a.mat, o + L.o_jp, 1.f);
    }
    if (w == 0 && !L.h) {
        c.Pjm = Pair<S>::raw(a.in, o + L.o_jm);
        c.rjm = Pair<float>::ld(a.mat, o + L.o_jm, 1.f);
        c.vyjm = Pair<S>::raw(a.in, F_VY * a.nnn + o + L.o_jm);
    }
}

// CPML memory-variable addresses of one cell (or -1 outside the slabs)
struct Slab {
    int sy, sz;   // y, z slab coordinates
};

// memory variable of the velocity set (mem_in -> mem_out) or the pressure
// set (memp, in place) at cell (i, j, k): x slab index sx, y index sy,
// z index sz; the value 0 where the cell is outside that slab
template <typename S>
__device__ __forceinline__ unsigned mx_off(const FusedArgs<S> &a, int sx, unsigned plane,
                                           unsigned ojk)
{
    return (unsigned)sx * plane + ojk;
}
template <typename S>
__device__ __forceinline__ unsigned my_off(const FusedArgs<S> &a, int i, int sy, int k)
{
    return a.my + ((unsigned)i * 2u * a.np + sy) * a.pz + k;
}
template <typename S>
__device__ __forceinline__ unsigned mz_off(const FusedArgs<S> &a, int i, int j, int sz)
{
    return a.mz + ((unsigned)i * a.ny2 + j) * 2u * a.np + sz;
}

// velocity update of one cell with its optional CPML memory variable:
// m is read from `min` at offset `off` and written to `mout` (velocity set)
template <typename S>
__device__ __forceinline__ float vel_cell(float v_old, float p0, float p1, float r0, float r1,
                                          float ih, float two_dt, bool slab, float bh, float ah,
                                          float ikh, const S *min_, S *mout, unsigned off,
                                          float sM, float isM, unsigned salt, int sr, bool store)
{
    float m = 0.f;
    if (slab) m = Store<S>::ld(min_, off, sM);
    const float v = vel_update(v_old, p0, p1, r0, r1, ih, two_dt, slab, bh, ah, ikh, m);
    if (slab && store) Store<S>::st(mout, off, m, isM, salt, sr);
    return v;
}

template <typename S, int TJ>
__global__ void __launch_bounds__(64 * TJ)
a3d_pair_step(FusedArgs<S> a)
{
    // new vy of each wave's upper row, by cell 2m+e
    __shared__ float s_vy[2][TJ][64];

    // ---------------- receiver gather blocks (previous step) -------------
    if ((int)blockIdx.x >= a.nblocks) {
        if (a.gather_row < 0) return;
        const int r = ((int)blockIdx.x - a.nblocks) * (64 * TJ) + threadIdx.y * 64 + threadIdx.x;
        if (r < a.nrec) {
            const int cp = a.rcomp[r];
            const unsigned q = cp == SM_VX ? 1u : cp == SM_VY ? 2u : cp == SM_VZ ? 3u : 0u;
            const float sc = cp == SM_P ? a.sP : a.sV;
            a.trace[(size_t)a.gather_row * a.nrec + r] =
                Store<S>::ld(a.in, q * a.nnn + (unsigned)a.ridx[r], sc);
        }
        return;
    }

    const int lane = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const unsigned plane = (unsigned)a.ny2 * a.pz;
    const int nx2 = a.nx2, ny2 = a.ny2, nz2 = a.nz2, np = a.np;
    const float *ty = a.tab + a.toff_y, *tz = a.tab + a.toff_z;

    for (int item = blockIdx.x; item < a.nitems; item += a.nblocks) {
        const int logical = (item % 8) * (a.nitems / 8) + item / 8;
        __syncthreads();   // the LDS of the previous item is no longer read
        if (logical >= a.nreal) continue;   // uniform per block
        const int chunk = logical / a.ntiles;
        const int tile = logical - chunk * a.ntiles;
        const int tjb = tile / a.ntk;
        const int tk = tile - tjb * a.ntk;

        PLane L;
        L.h = lane >> 5;
        const int m = lane & 31;
        L.j = tjb * 2 * TJ + 2 * w + L.h;
        L.k0 = tk * 62 + 2 * m;
        const bool jv = L.j < ny2;
        L.own0 = jv && L.k0 < nz2 && (m > 0 || tk == 0);
        L.own1 = jv && L.k0 + 1 < nz2 && m < 31;
        const int jc = min(L.j, ny2 - 1);
        const int kp = min(L.k0, nz2 - 2);    // pair start, clamped to the last pair
        L.o = (unsigned)jc * a.pz + kp;
        L.o_jp = (unsigned)min(jc + 1, ny2 - 1) * a.pz + kp;
        L.o_jm = (unsigned)max(jc - 1, 0) * a.pz + kp;
        // y CPML of this row (velocity: rows j < ny2-1; pressure: all rows)
        // and of the row below (wave 0's halo row)
        const int syv = jv && L.j < ny2 - 1 ? slab_vel(L.j, ny2, np) : -1;
        const int syp = jv ? slab_pre(L.j, ny2, np) : -1;
        const int syh = jv && L.j >= 1 ? slab_vel(L.j - 1, ny2, np) : -1;
        // z CPML of the two cells
        const int k1 = L.k0 + 1;
        const int szv0 = L.k0 < nz2 - 1 ? slab_vel(L.k0, nz2, np) : -1;
        const int szv1 = k1 < nz2 - 1 ? slab_vel(k1, nz2, np) : -1;
        const int szp0 = L.k0 < nz2 ? slab_pre(L.k0, nz2, np) : -1;
        const int szp1 = k1 < nz2 ? slab_pre(k1, nz2, np) : -1;
        // does this wave have any y- or z-slab cell (uniform)
        const bool yslab = __any(syv >= 0 || syp >= 0 || (w == 0 && syh >= 0));
        const bool zslab = __any(szv0 >= 0 || szv1 >= 0 || szp0 >= 0 || szp1 >= 0);

        const int i0 = chunk * a.chunk_len;
        const int i1 = min(nx2, i0 + a.chunk_len);
        if (i0 >= nx2) continue;   // uniform; the host never makes empty chunks
        const int ist = i0 > 0 ? i0 - 1 : 0;

        float2 P_c = Pair<S>::ld(a.in, (unsigned)ist * plane + L.o, a.sP);
        float2 r_c = Pair<float>::ld(a.mat, (unsigned)ist * plane + L.o, 1.f);
        float2 vx_prev = make_float2(0.f, 0.f);
        const float2 z2 = make_float2(0.f, 0.f);
        PCol<S> A, B;
        A.Pj = A.Pjm = A.vyjm = typename Pair<S>::Raw{};
        A.rj = A.rjm = z2;
        B = A;
        pload<S, TJ>(a, L, w, ist, plane, A);

        // one plane: prefetch plane i+1 into nxt, update plane i from cur
        auto plane_step = [&](int i, PCol<S> &cur, PCol<S> &nxt) {
            if (i + 1 < i1) pload<S, TJ>(a, L, w, i + 1, plane, nxt);
            const XPlane x = xplane(a.xpl, i);
            const unsigned o = (unsigned)i * plane + L.o;   // first cell of the pair
            const float2 Pn_ = Pair<S>::unpack(cur.Pn, a.sP);
            const float2 vx_ = Pair<S>::unpack(cur.vx, a.sV);

            // ---- vx (a_iso_3d.c:605-619); last plane: dP = 0 ----
            float2 vxn;
            {
                const bool sl = x.sxv >= 0;
                const unsigned mo = sl ? mx_off(a, x.sxv, plane, L.o) : 0u;
                const bool st = i >= i0;
                vxn.x = vel_cell<S>(vx_.x, P_c.x, Pn_.x, r_c.x, cur.rn.x, a.idx, a.two_dt, sl,
                                    x.bh, x.ah, x.ikh, a.mem_in, a.mem_out, mo, a.sMv, a.isMv,
                                    a.salt ^ SALT_MV, a.sr, st && L.own0);
                vxn.y = vel_cell<S>(vx_.y, P_c.y, Pn_.y, r_c.y, cur.rn.y, a.idx, a.two_dt, sl,
                                    x.bh, x.ah, x.ikh, a.mem_in, a.mem_out, mo + 1, a.sMv, a.isMv,
                                    a.salt ^ SALT_MV, a.sr, st && L.own1);
            }
            if (i < i0) {   // chunk start: only vx of the plane below is needed
                vx_prev = vxn;
                P_c = Pn_;
                r_c = cur.rn;
                return;
            }

            // ---- neighbours ----
            // row above: the upper half for the lower half, loaded for the upper half
            float2 Pj, rj;
            {
                const float2 uP = make_float2(__shfl_down(P_c.x, 32), __shfl_down(P_c.y, 32));
                const float2 ur = make_float2(__shfl_down(r_c.x, 32), __shfl_down(r_c.y, 32));
                Pj = L.h ? Pair<S>::unpack(cur.Pj, a.sP) : uP;
                rj = L.h ? cur.rj : ur;
            }
            // z+1: the second cell of the pair, or the next lane's first cell;
            // the last cell of the grid has no update (dP = 0)
            float2 Pk, rk;
            Pk.x = P_c.y;
            rk.x = r_c.y;
            Pk.y = __shfl_down(P_c.x, 1);
            rk.y = __shfl_down(r_c.x, 1);
            if (k1 >= nz2 - 1) { Pk.y = P_c.y; rk.y = r_c.y; }

            // ---- vy, vz (a_iso_3d.c:621-651) ----
            float2 vyn, vzn;
            const float2 vy_ = Pair<S>::unpack(cur.vy, a.sV);
            const float2 vz_ = Pair<S>::unpack(cur.vz, a.sV);
            if (yslab) {
                const bool sl = syv >= 0;
                float yb = 0.f, ya = 0.f, yk = 1.f;
                if (sl) {
                    const int jn = min(L.j, ny2 - 2);
                    yb = ty[4 * ny2 + jn]; ya = ty[3 * ny2 + jn]; yk = ty[5 * ny2 + jn];
                }
                const unsigned mo = sl ? my_off(a, i, syv, kp) : 0u;
                vyn.x = vel_cell<S>(vy_.x, P_c.x, Pj.x, r_c.x, rj.x, a.idy, a.two_dt, sl, yb,
                                    ya, yk, a.mem_in, a.mem_out, mo, a.sMv, a.isMv,
                                    a.salt ^ SALT_MV, a.sr, L.own0);
                vyn.y = vel_cell<S>(vy_.y, P_c.y, Pj.y, r_c.y, rj.y, a.idy, a.two_dt, sl, yb,
                                    ya, yk, a.mem_in, a.mem_out, mo + 1, a.sMv, a.isMv,
                                    a.salt ^ SALT_MV, a.sr, L.own1);
            } else {
                float dm = 0.f;
                vyn.x = vel_update(vy_.x, P_c.x, Pj.x, r_c.x, rj.x, a.idy, a.two_dt, false, 0.f,
                                   0.f, 1.f, dm);
                vyn.y = vel_update(vy_.y, P_c.y, Pj.y, r_c.y, rj.y, a.idy, a.two_dt, false, 0.f,
                                   0.f, 1.f, dm);
            }
            if (zslab) {
                const bool s0 = szv0 >= 0, s1 = szv1 >= 0;
                const int kn0 = min(L.k0, nz2 - 2), kn1 = min(k1, nz2 - 2);
                const float b0 = s0 ? tz[4 * nz2 + kn0] : 0.f, a0 = s0 ? tz[3 * nz2 + kn0] : 0.f;
                const float c0 = s0 ? tz[5 * nz2 + kn0] : 1.f;
                const float b1 = s1 ? tz[4 * nz2 + kn1] : 0.f, a1 = s1 ? tz[3 * nz2 + kn1] : 0.f;
                const float c1 = s1 ? tz[5 * nz2 + kn1] : 1.f;
                vzn.x = vel_cell<S>(vz_.x, P_c.x, Pk.x, r_c.x, rk.x, a.idz, a.two_dt, s0, b0, a0,
                                    c0, a.mem_in, a.mem_out, s0 ? mz_off(a, i, jc, szv0) : 0u,
                                    a.sMv, a.isMv, a.salt ^ SALT_MV, a.sr, L.own0);
                vzn.y = vel_cell<S>(vz_.y, P_c.y, Pk.y, r_c.y, rk.y, a.idz, a.two_dt, s1, b1, a1,
                                    c1, a.mem_in, a.mem_out, s1 ? mz_off(a, i, jc, szv1) : 0u,
                                    a.sMv, a.isMv, a.salt ^ SALT_MV, a.sr, L.own1);
            } else {
                float dm = 0.f;
                vzn.x = vel_update(vz_.x, P_c.x, Pk.x, r_c.x, rk.x, a.idz, a.two_dt, false, 0.f,
                                   0.f, 1.f, dm);
                vzn.y = vel_update(vz_.y, P_c.y, Pk.y, r_c.y, rk.y, a.idz, a.two_dt, false, 0.f,
                                   0.f, 1.f, dm);
            }
            // store the new velocities (pairs where both cells are owned)
            if (L.own0 && L.own1) {
                Pair<S>::st(a.out, F_VX * a.nnn + o, vxn, a.isV, a.salt, a.sr);
                Pair<S>::st(a.out, F_VY * a.nnn + o, vyn, a.isV, a.salt, a.sr);
                Pair<S>::st(a.out, F_VZ * a.nnn + o, vzn, a.isV, a.salt, a.sr);
            } else if (L.own0) {
                Store<S>::st(a.out, F_VX * a.nnn + o, vxn.x, a.isV, a.salt, a.sr);
                Store<S>::st(a.out, F_VY * a.nnn + o, vyn.x, a.isV, a.salt, a.sr);
                Store<S>::st(a.out, F_VZ * a.nnn + o, vzn.x, a.isV, a.salt, a.sr);
            } else if (L.own1) {
                Store<S>::st(a.out, F_VX * a.nnn + o + 1, vxn.y, a.isV, a.salt, a.sr);
                Store<S>::st(a.out, F_VY * a.nnn + o + 1, vyn.y, a.isV, a.salt, a.sr);
                Store<S>::st(a.out, F_VZ * a.nnn + o + 1, vzn.y, a.isV, a.salt, a.sr);
            }

            // ---- vy of the row below the lower half of wave 0 (not stored) ----
            float2 vy_halo = z2;
            if (w == 0) {
                const float2 vyjm = Pair<S>::unpack(cur.vyjm, a.sV);
                const float2 Pjm = Pair<S>::unpack(cur.Pjm, a.sP);
                const bool sl = syh >= 0 && !L.h;
                float hb = 0.f, ha = 0.f, hk = 1.f;
                if (sl) {
                    const int jh = max(min(L.j - 1, ny2 - 2), 0);
                    hb = ty[4 * ny2 + jh]; ha = ty[3 * ny2 + jh]; hk = ty[5 * ny2 + jh];
                }
                const unsigned mo = sl ? my_off(a, i, syh, kp) : 0u;
                vy_halo.x = vel_cell<S>(vyjm.x, Pjm.x, P_c.x, cur.rjm.x, r_c.x, a.idy,
                                        a.two_dt, sl, hb, ha, hk, a.mem_in, a.mem_out, mo, a.sMv,
                                        a.isMv, 0u, 0, false);
                vy_halo.y = vel_cell<S>(vyjm.y, Pjm.y, P_c.y, cur.rjm.y, r_c.y, a.idy,
                                        a.two_dt, sl, hb, ha, hk, a.mem_in, a.mem_out, mo + 1,
                                        a.sMv, a.isMv, 0u, 0, false);
            }
            if (L.h) {
                s_vy[i & 1][w][2 * (lane & 31)] = vyn.x;
                s_vy[i & 1][w][2 * (lane & 31) + 1] = vyn.y;
            }
            __syncthreads();
            float2 vy_m;
            {
                const float2 lo = make_float2(__shfl_up(vyn.x, 32), __shfl_up(vyn.y, 32));
                float2 below = vy_halo;
                if (w > 0) below = make_float2(s_vy[i & 1][w - 1][2 * m], s_vy[i & 1][w - 1][2 * m + 1]);
                vy_m = L.h ? lo : below;
            }
            float2 vz_m;
            vz_m.y = vzn.x;
            vz_m.x = __shfl_up(vzn.y, 1);   // lanes 0 and 32 hold the left halo cell

            // ---- pressure (a_iso_3d.c:659-683) ----
            float2 Pn = P_c;
            const bool upd0 = L.own0 && i >= 1 && L.j >= 1 && L.k0 >= 1;
            const bool upd1 = L.own1 && i >= 1 && L.j >= 1;
            if (upd0 || upd1) {
                float dvdx0 = (vxn.x - vx_prev.x) * a.idx, dvdx1 = (vxn.y - vx_prev.y) * a.idx;
                float dvdy0 = (vyn.x - vy_m.x) * a.idy, dvdy1 = (vyn.y - vy_m.y) * a.idy;
                float dvdz0 = (vzn.x - vz_m.x) * a.idz, dvdz1 = (vzn.y - vz_m.y) * a.idz;
                if (x.sxp >= 0) {
                    const unsigned mo = mx_off(a, x.sxp, plane, L.o);
                    const float2 mv = Pair<S>::ld(a.memp, mo, a.sMp);
                    const float m0 = x.b * mv.x + x.a * dvdx0, m1 = x.b * mv.y + x.a * dvdx1;
                    if (upd0) Store<S>::st(a.memp, mo, m0, a.isMp, a.salt ^ SALT_MP, a.sr);
                    if (upd1) Store<S>::st(a.memp, mo + 1, m1, a.isMp, a.salt ^ SALT_MP, a.sr);
                    dvdx0 = dvdx0 * x.ik + m0;
                    dvdx1 = dvdx1 * x.ik + m1;
                }
                if (syp >= 0) {
                    const int jc2 = min(L.j, ny2 - 1);
                    const float yb = ty[ny2 + jc2], ya = ty[jc2], yk = ty[2 * ny2 + jc2];
                    const unsigned mo = my_off(a, i, syp, kp);
                    const float2 mv = Pair<S>::ld(a.memp, mo, a.sMp);
                    const float m0 = yb * mv.x + ya * dvdy0, m1 = yb * mv.y + ya * dvdy1;
                    if (upd0) Store<S>::st(a.memp, mo, m0, a.isMp, a.salt ^ SALT_MP, a.sr);
                    if (upd1) Store<S>::st(a.memp, mo + 1, m1, a.isMp, a.salt ^ SALT_MP, a.sr);
                    dvdy0 = dvdy0 * yk + m0;
                    dvdy1 = dvdy1 * yk + m1;
                }
                if (szp0 >= 0) {
                    const int kc = min(L.k0, nz2 - 1);
                    const unsigned mo = mz_off(a, i, jc, szp0);
                    const float m0 = tz[nz2 + kc] * Store<S>::ld(a.memp, mo, a.sMp) + tz[kc] * dvdz0;
                    if (upd0) Store<S>::st(a.memp, mo, m0, a.isMp, a.salt ^ SALT_MP, a.sr);
                    dvdz0 = dvdz0 * tz[2 * nz2 + kc] + m0;
                }
                if (szp1 >= 0) {
                    const int kc = min(k1, nz2 - 1);
                    const unsigned mo = mz_off(a, i, jc, szp1);
                    const float m1 = tz[nz2 + kc] * Store<S>::ld(a.memp, mo, a.sMp) + tz[kc] * dvdz1;
                    if (upd1) Store<S>::st(a.memp, mo, m1, a.isMp, a.salt ^ SALT_MP, a.sr);
                    dvdz1 = dvdz1 * tz[2 * nz2 + kc] + m1;
                }
                if (upd0) Pn.x = P_c.x - cur.K.x * (dvdx0 + dvdy0 + dvdz0);
                if (upd1) Pn.y = P_c.y - cur.K.y * (dvdx1 + dvdy1 + dvdz1);
            }
            // sources of this plane (a_iso_3d.c:689-693), in source order
            for (int q = x.q0; q < x.q1; ++q) {
                const unsigned c = (unsigned)a.src[q];
                if (a.step < a.src[2 * a.nsrc + q]) {
                    const float inc = a.src_inc[a.src[a.nsrc + q] + a.step];
                    if (c == o && L.own0) Pn.x += inc;
                    if (c == o + 1 && L.own1) Pn.y += inc;
                }
            }
            if (L.own0 && L.own1) Pair<S>::st(a.out, o, Pn, a.isP, a.salt, a.sr);
            else if (L.own0) Store<S>::st(a.out, o, Pn.x, a.isP, a.salt, a.sr);
            else if (L.own1) Store<S>::st(a.out, o + 1, Pn.y, a.isP, a.salt, a.sr);

            vx_prev = vxn;
            P_c = Pn_;
            r_c = cur.rn;
        };
        for (int i = ist; i < i1; i += 2) {
            plane_step(i, A, B);
            if (i + 1 >= i1) break;
            plane_step(i + 1, B, A);
        }
    }
}

template <typename S, int TJ>
const void *pair_ptr()
{
    return reinterpret_cast<const void *>(&a3d_pair_step<S, TJ>);
}

}  // namespace

template <typename S>
int pair_max_blocks_per_cu(int tj, int *blocks)
{
    const void *f = nullptr;
    switch (tj) {
    case 1: f = pair_ptr<S, 1>(); break;
    case 2: f = pair_ptr<S, 2>(); break;
    case 4: f = pair_ptr<S, 4>(); break;
    case 8: f = pair_ptr<S, 8>(); break;
    default: return SM_ERR_ARG;
    }
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64 * tj, 0) != hipSuccess)
        return SM_ERR_HIP;
    *blocks = nb;
    return SM_OK;
}

template <typename S>
int pair_launch(const FusedArgs<S> &a, int tj, hipStream_t st)
{
    const int blocks = a.nblocks + a.gather_blocks;
    switch (tj) {
    case 1: a3d_pair_step<S, 1><<<blocks, dim3(64, 1), 0, st>>>(a); break;
    case 2: a3d_pair_step<S, 2><<<blocks, dim3(64, 2), 0, st>>>(a); break;
    case 4: a3d_pair_step<S, 4><<<blocks, dim3(64, 4), 0, st>>>(a); break;
    case 8: a3d_pair_step<S, 8><<<blocks, dim3(64, 8), 0, st>>>(a); break;
    default: return SM_ERR_ARG;
    }
    return hipGetLastError() == hipSuccess ? SM_OK : SM_ERR_HIP;
}

template int pair_launch<float>(const FusedArgs<float> &, int, hipStream_t);
template int pair_launch<__hip_bfloat16>(const FusedArgs<__hip_bfloat16> &, int, hipStream_t);
template int pair_launch<__half>(const FusedArgs<__half> &, int, hipStream_t);
template int pair_max_blocks_per_cu<float>(int, int *);
template int pair_max_blocks_per_cu<__hip_bfloat16>(int, int *);
template int pair_max_blocks_per_cu<__half>(int, int *);

}  // namespace sm
