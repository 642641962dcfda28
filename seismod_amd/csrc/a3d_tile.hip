// a3d_tile.hip -- single-pass 3D acoustic CPML step with LDS-staged tiles,
// FP32 arithmetic, FP32 / BF16 / FP16 storage.
//
// The same time step as a3d_fused_step (a_iso_3d.c:599-693 plus the
// receivers of the previous step) with the same per-cell expressions, so
// results are bit-identical to it; what differs is how values reach the
// lanes.
//
//   - a block owns TJ rows (one wave per row along y) by 64 cells (lanes
//     along z, tiles aligned to 64 cells) and marches along x over a chunk of
//     planes, as the other variants do (persistent grid, XCD-contiguous
//     item order, chunks with i0 > 0 recompute vx of plane i0-1).
//   - for every plane the block copies what it reads into LDS: P and rho
//     with one row and one 16-byte chunk of halo on each side, old vy with
//     the row below, old vz with the chunk to the left, old vx and K*dt.
//     Each lane copies 16-byte chunks (global_load_dwordx4, ds_write_b128),
//     so every row of every field crosses the memory pipeline once per
//     plane, in whole chunks, and all neighbour values come from LDS.  The
//     other variants issue about twice as many vector-memory instructions
//     per cell, most of them 4-byte-per-lane loads of rows other waves also
//     load, which made their speed depend on L2 reuse.
//   - three plane slots rotate: while plane i is updated from slots i and
//     i+1, the chunks of plane i+2 (loaded during the previous plane) are
//     written to the third slot and the loads of plane i+3 are issued.  One
//     barrier per plane.
//   - each wave recomputes new vy of the row below its own and lane 0 new
//     vz of the cell left of the tile, with the expressions their owners use,
//     so no values pass between waves within a plane.
//   - CPML memory variables exist only in the slabs; each lane loads its own
//     one plane ahead.
//
// Requirements: the device row pitch is a multiple of 8 cells (16 bytes of
// 16-bit storage), so chunks never straddle a row end.
#include <cstdint>

#include "a3d_common.hpp"
#include "a3d_fused.hpp"
#include "a3d_store.hpp"
#include "seismod_hip.h"

namespace sm {

namespace {

// Tile fields of one plane slot, in LDS order.
enum { T_P = 0, T_R = 1, T_VY = 2, T_VZ = 3, T_VX = 4, T_K = 5, T_NF = 6 };

// Geometry of the staged tiles for storage type S and TJ rows per block.
// Rows are relative to j0 (the block's first row), cells to k0 (first cell).
template <typename S, int TJ>
struct TileShape {
    static constexpr int ES = (int)sizeof(S);
    static constexpr int esz(int f) { return (f == T_R || f == T_K) ? 4 : ES; }
    static constexpr int cpc(int f) { return 16 / esz(f); }   // cells per 16-byte chunk
    static constexpr int rows(int f) { return f <= T_R ? TJ + 2 : f == T_VY ? TJ + 1 : TJ; }
    static constexpr int row0(int f) { return f <= T_VY ? -1 : 0; }
    static constexpr int c0(int f) { return (f <= T_R || f == T_VZ) ? -cpc(f) : 0; }
    static constexpr int cells(int f)
    {
        return f <= T_R ? 64 + 2 * cpc(f) : f == T_VZ ? 64 + cpc(f) : 64;
    }
    static constexpr int row_bytes(int f) { return cells(f) * esz(f); }
    static constexpr int cpr(int f) { return row_bytes(f) / 16; }   // chunks per row
    static constexpr int chunks(int f) { return rows(f) * cpr(f); }
    static constexpr int chunk0(int f) { return f == 0 ? 0 : chunk0(f - 1) + chunks(f - 1); }
    static constexpr int off(int f) { return f == 0 ? 0 : off(f - 1) + rows(f - 1) * row_bytes(f - 1); }
    static constexpr int total_chunks = chunk0(T_NF - 1) + chunks(T_NF - 1);
    static constexpr int slot_bytes = off(T_NF - 1) + rows(T_NF - 1) * row_bytes(T_NF - 1);
    static constexpr int NT = 64 * TJ;                                // threads per block
    static constexpr int NR = (total_chunks + NT - 1) / NT;          // chunks per thread
};

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// One staged 16-byte chunk of a lane: its address in plane 0, bytes per
// plane, and its byte offset within a slot (base == nullptr: no chunk).
struct Chunk {
    const char *base;
    unsigned pstride;
    unsigned lds;
};

// Per-item constants of one lane (the cell (j, kk) it owns, clamped copies,
// slab coordinates, CPML coefficients and LDS offsets of its reads).
struct TLane {
    int j, kk, kc;
    bool own;
    unsigned o_jk;                       // in-plane device offset of the cell
    int syv, syp, syh, szv, szp, szh;    // slab coordinates (-1: not in a slab)
    unsigned mz_v, mz_p, mz_h;           // z-slab offsets within a plane
    float ybh, yah, yikh, yb, ya, yik, hbh, hah, hikh;
    float zbh, zah, zikh, zb, za, zik;
    // tile coordinates of the lane's reads: rows (relative to j0, wave-
    // uniform) of the cell, j+1 and j-1; columns (relative to k0) of the
    // cell and k+1 (per lane) and of k0-1 (uniform)
    int tc, tjp, tjm;
    int cc, ckp, ckm;
};

// LDS byte offset within a slot of field F at tile row t, column c (the
// field is a template argument so the geometry folds to constants)
template <typename TS, int F>
__device__ __forceinline__ unsigned tof(int t, int c)
{
    constexpr int O = TS::off(F), R0 = TS::row0(F), RB = TS::row_bytes(F), C0 = TS::c0(F),
                  E = TS::esz(F);
    return (unsigned)(O + (t - R0) * RB + (c - C0) * E);
}

// The chunk c of a slot if it belongs to field F: its address in plane 0,
// bytes per plane and LDS offset.  Tile rows are clamped to the grid;
// chunks past the row end (no cell of the grid) and before k = 0 (halo of
// the first tile, never read) are clamped into the row.
template <typename S, typename TS, int F>
__device__ __forceinline__ void chunk_of(const FusedArgs<S> &a, int c, int j0, int k0,
                                         unsigned plane, Chunk &ck)
{
    constexpr int C0 = TS::chunk0(F), NC = TS::chunks(F), CPR = TS::cpr(F), CPC = TS::cpc(F),
                  R0 = TS::row0(F), Z0 = TS::c0(F), E = TS::esz(F), O = TS::off(F),
                  RB = TS::row_bytes(F);
    if (c < C0 || c >= C0 + NC) return;
    const int q = c - C0;
    const int row = q / CPR, cc = q - row * CPR;
    const int jr = min(max(j0 + R0 + row, 0), a.ny2 - 1);
    int cs = k0 + Z0 + cc * CPC;
    cs = cs < 0 ? 0 : cs >= a.pz ? a.pz - CPC : cs;
    const char *fb;
    if constexpr (F == T_R) fb = (const char *)a.mat;
    else if constexpr (F == T_K) fb = (const char *)(a.mat + a.nnn);
    else {
        constexpr unsigned qf = F == T_P ? 0u : F == T_VX ? F_VX : F == T_VY ? F_VY : F_VZ;
        fb = (const char *)(a.in + (size_t)qf * a.nnn);
    }
    ck.base = fb + ((size_t)jr * a.pz + cs) * E;
    ck.pstride = plane * (unsigned)E;
    ck.lds = (unsigned)(O + row * RB + cc * 16);
}

// CPML memory variables one lane reads for one plane (0 outside the slabs)
struct Psi {
    float mPx, mPy, mPz;   // velocity set at the own cell
    float mPym;            // velocity set (y) at row j-1
    float mPzm;            // velocity set (z) at k0-1 (lane 0)
    float mvx, mvy, mvz;   // pressure set at the own cell
};

template <typename S>
__device__ __forceinline__ float lds_state(const char *slot, unsigned off, float scale)
{
    return Store<S>::ld((const S *)(slot + off), 0u, scale);
}
__device__ __forceinline__ float lds_f32(const char *slot, unsigned off)
{
    return *(const float *)(slot + off);
}

template <typename S>
__device__ __forceinline__ void load_psi(const FusedArgs<S> &a, const TLane &L, int i,
                                         unsigned plane, Psi &m)
{
    const XPlane x = xplane(a.xpl, i);
    const unsigned np2 = 2u * a.np;
    const unsigned yrow = (unsigned)i * np2;
    const unsigned zpl = a.mz + (unsigned)i * a.ny2 * np2;
    m.mPx = x.sxv >= 0 ? Store<S>::ld(a.mem_in, (unsigned)x.sxv * plane + L.o_jk, a.sMv) : 0.f;
    m.mPy = L.syv >= 0 ? Store<S>::ld(a.mem_in, a.my + (yrow + L.syv) * a.pz + L.kc, a.sMv) : 0.f;
    m.mPz = L.szv >= 0 ? Store<S>::ld(a.mem_in, zpl + L.mz_v, a.sMv) : 0.f;
    m.mPym = L.syh >= 0 ? Store<S>::ld(a.mem_in, a.my + (yrow + L.syh) * a.pz + L.kc, a.sMv) : 0.f;
    m.mPzm = L.szh >= 0 ? Store<S>::ld(a.mem_in, zpl + L.mz_h, a.sMv) : 0.f;
    m.mvx = x.sxp >= 0 ? Store<S>::ld(a.memp, (unsigned)x.sxp * plane + L.o_jk, a.sMp) : 0.f;
    m.mvy = L.syp >= 0 ? Store<S>::ld(a.memp, a.my + (yrow + L.syp) * a.pz + L.kc, a.sMp) : 0.f;
    m.mvz = L.szp >= 0 ? Store<S>::ld(a.memp, zpl + L.mz_p, a.sMp) : 0.f;
}

// issue the chunk loads of plane p (clamped to the last plane)
template <int NR>
__device__ __forceinline__ void stage_issue(const Chunk (&ck)[NR], int p, u32x4 (&v)[NR])
{
#pragma unroll
    for (int r = 0; r < NR; ++r)
        if (ck[r].base) v[r] = *(const u32x4 *)(ck[r].base + (size_t)p * ck[r].pstride);
}

template <int NR>
__device__ __forceinline__ void stage_write(const Chunk (&ck)[NR], char *slot, const u32x4 (&v)[NR])
{
#pragma unroll
    for (int r = 0; r < NR; ++r)
        if (ck[r].base) *(u32x4 *)(slot + ck[r].lds) = v[r];
}

// update plane i from slot cs (plane i) and ns (plane i+1); same expressions
// and operand order as the other variants' plane update
template <typename S, typename TS>
__device__ __forceinline__ void tile_plane(const FusedArgs<S> &a, const TLane &L, int i, int i0,
                                           unsigned plane, const char *cs, const char *ns,
                                           const Psi &m, const Halo &hz, float &vx_prev)
{
    const unsigned p_c = tof<TS, T_P>(L.tc, L.cc), r_c_ = tof<TS, T_R>(L.tc, L.cc);
    const XPlane x = xplane(a.xpl, i);
    const unsigned o = (unsigned)i * plane + L.o_jk;
    const unsigned np2 = 2u * a.np;
    const float P_c = lds_state<S>(cs, p_c, a.sP);
    const float r_c = lds_f32(cs, r_c_);

    // vx at (i, j, k), a_iso_3d.c:605-619 (the next plane clamped to the last)
    float mx = m.mPx;
    const float vxn = vel_update(lds_state<S>(cs, tof<TS, T_VX>(L.tc, L.cc), a.sV), P_c,
                                 lds_state<S>(ns, p_c, a.sP), r_c, lds_f32(ns, r_c_), a.idx, a.two_dt, x.sxv >= 0, x.bh, x.ah,
                                 x.ikh, mx);
    if (i < i0) {   // chunk start: only vx of the plane below is needed
        vx_prev = vxn;
        return;
    }
    // vy, vz at (i, j, k), a_iso_3d.c:621-651
    float my = m.mPy, mz = m.mPz;
    const float vyn = vel_update(lds_state<S>(cs, tof<TS, T_VY>(L.tc, L.cc), a.sV), P_c,
                                 lds_state<S>(cs, tof<TS, T_P>(L.tjp, L.cc), a.sP), r_c,
                                 lds_f32(cs, tof<TS, T_R>(L.tjp, L.cc)), a.idy, a.two_dt, L.syv >= 0, L.ybh, L.yah,
                                 L.yikh, my);
    const float vzn = vel_update(lds_state<S>(cs, tof<TS, T_VZ>(L.tc, L.cc), a.sV), P_c,
                                 lds_state<S>(cs, tof<TS, T_P>(L.tc, L.ckp), a.sP), r_c,
                                 lds_f32(cs, tof<TS, T_R>(L.tc, L.ckp)), a.idz, a.two_dt, L.szv >= 0, L.zbh, L.zah,
                                 L.zikh, mz);
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
    // new vy of row j-1 and new vz of k0-1 (owned by other waves / tiles;
    // recomputed here, not stored)
    float mh = m.mPym;
    const float vy_m = vel_update(lds_state<S>(cs, tof<TS, T_VY>(L.tjm, L.cc), a.sV),
                                  lds_state<S>(cs, tof<TS, T_P>(L.tjm, L.cc), a.sP), P_c,
                                  lds_f32(cs, tof<TS, T_R>(L.tjm, L.cc)), r_c, a.idy, a.two_dt, L.syh >= 0, L.hbh, L.hah,
                                  L.hikh, mh);
    float vz_m = __shfl_up(vzn, 1);
    {
        float mzh = m.mPzm;
        const float vz_h = vel_update(lds_state<S>(cs, tof<TS, T_VZ>(L.tc, L.ckm), a.sV),
                                      lds_state<S>(cs, tof<TS, T_P>(L.tc, L.ckm), a.sP), P_c,
                                      lds_f32(cs, tof<TS, T_R>(L.tc, L.ckm)), r_c, a.idz, a.two_dt, L.szh >= 0, hz.bh,
                                      hz.ah, hz.ikh, mzh);
        if (threadIdx.x == 0) vz_m = vz_h;
    }

    // pressure at (i, j, k), a_iso_3d.c:659-683
    if (L.own) {
        float Pn = P_c;
        if (i >= 1 && L.j >= 1 && L.kk >= 1) {
            float dvdx = (vxn - vx_prev) * a.idx;
            float dvdy = (vyn - vy_m) * a.idy;
            float dvdz = (vzn - vz_m) * a.idz;
            if (x.sxp >= 0) {
                const float mm = x.b * m.mvx + x.a * dvdx;
                Store<S>::st(a.memp, (unsigned)x.sxp * plane + L.o_jk, mm, a.isMp,
                             a.salt ^ SALT_MP, a.sr);
                dvdx = dvdx * x.ik + mm;
            }
            if (L.syp >= 0) {
                const float mm = L.yb * m.mvy + L.ya * dvdy;
                Store<S>::st(a.memp, a.my + ((unsigned)i * np2 + L.syp) * a.pz + L.kc, mm, a.isMp,
                             a.salt ^ SALT_MP, a.sr);
                dvdy = dvdy * L.yik + mm;
            }
            if (L.szp >= 0) {
                const float mm = L.zb * m.mvz + L.za * dvdz;
                Store<S>::st(a.memp, a.mz + (unsigned)i * a.ny2 * np2 + L.mz_p, mm, a.isMp,
                             a.salt ^ SALT_MP, a.sr);
                dvdz = dvdz * L.zik + mm;
            }
            Pn = P_c - lds_f32(cs, tof<TS, T_K>(L.tc, L.cc)) * (dvdx + dvdy + dvdz);
        }
        // sources of this plane (a_iso_3d.c:689-693), in source order
        for (int q = x.q0; q < x.q1; ++q)
            if ((unsigned)a.src[q] == o && a.step < a.src[2 * a.nsrc + q])
                Pn += a.src_inc[a.src[a.nsrc + q] + a.step];
        Store<S>::st(a.out, o, Pn, a.isP, a.salt, a.sr);
    }
    vx_prev = vxn;
}

}  // namespace

template <typename S, int TJ>
__global__ void __launch_bounds__(64 * TJ)
a3d_tile_step(FusedArgs<S> a)
{
    using TS = TileShape<S, TJ>;
    constexpr int NR = TS::NR;
    __shared__ __attribute__((aligned(16))) char s_tile[3 * TS::slot_bytes];

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
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.y);   // one wave per y row
    const int tid = w * 64 + lane;
    const unsigned plane = (unsigned)a.ny2 * a.pz;
    const unsigned np2 = 2u * a.np;

    for (int item = blockIdx.x; item < a.nitems; item += a.nblocks) {
        const int logical = (item % 8) * (a.nitems / 8) + item / 8;
        __syncthreads();   // the LDS of the previous item is no longer read
        if (logical >= a.nreal) continue;   // uniform per block
        const int chunk = logical / a.ntiles;
        const int tile = logical - chunk * a.ntiles;
        const int tj = tile / a.ntk;
        const int tk = tile - tj * a.ntk;
        const int j0 = tj * TJ, k0 = tk * 64;

        // ---- the lane's cell; cells past the grid edge read the last cell
        // (their neighbour differences vanish) and are not owned
        TLane L;
        L.kk = k0 + lane;
        L.j = j0 + w;
        const bool jv = L.j < a.ny2, kv = L.kk < a.nz2;
        L.own = jv && kv;
        const int jc = min(L.j, a.ny2 - 1);
        L.kc = min(L.kk, a.nz2 - 1);
        const int jp = min(jc + 1, a.ny2 - 1), kp = min(L.kc + 1, a.nz2 - 1), jm = max(jc - 1, 0);
        const int km = max(k0 - 1, 0);
        L.o_jk = (unsigned)jc * a.pz + L.kc;
        L.syv = jv && L.j < a.ny2 - 1 ? slab_vel(L.j, a.ny2, a.np) : -1;
        L.syp = jv ? slab_pre(L.j, a.ny2, a.np) : -1;
        L.syh = jv && L.j >= 1 ? slab_vel(L.j - 1, a.ny2, a.np) : -1;
        L.szv = kv && L.kk < a.nz2 - 1 ? slab_vel(L.kk, a.nz2, a.np) : -1;
        L.szp = kv ? slab_pre(L.kk, a.nz2, a.np) : -1;
        const int hk = k0 - 1;   // the cell whose new vz lane 0 recomputes
        L.szh = lane == 0 && hk >= 0 && hk < a.nz2 - 1 ? slab_vel(hk, a.nz2, a.np) : -1;
        L.mz_v = (unsigned)jc * np2 + (L.szv < 0 ? 0 : L.szv);
        L.mz_p = (unsigned)jc * np2 + (L.szp < 0 ? 0 : L.szp);
        L.mz_h = (unsigned)jc * np2 + (L.szh < 0 ? 0 : L.szh);
        Halo hz{0.f, 1.f, 1.f};
        {
            // CPML tables a, b, 1/k, ah, bh, 1/kh along y (row and row below)
            // and z (the lane's cell and k0-1)
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
            const int khn = max(min(hk, nz2 - 2), 0);
            hz.bh = tz[4 * nz2 + khn]; hz.ah = tz[3 * nz2 + khn]; hz.ikh = tz[5 * nz2 + khn];
        }
        // ---- tile coordinates of the lane's reads
        L.tc = jc - j0; L.tjp = jp - j0; L.tjm = jm - j0;
        L.cc = L.kc - k0; L.ckp = kp - k0; L.ckm = km - k0;
        // ---- the 16-byte chunks this thread copies
        Chunk ck[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int c = r * TS::NT + tid;
            ck[r].base = nullptr;
            ck[r].pstride = 0;
            ck[r].lds = 0;
            chunk_of<S, TS, T_P>(a, c, j0, k0, plane, ck[r]);
            chunk_of<S, TS, T_R>(a, c, j0, k0, plane, ck[r]);
            chunk_of<S, TS, T_VY>(a, c, j0, k0, plane, ck[r]);
            chunk_of<S, TS, T_VZ>(a, c, j0, k0, plane, ck[r]);
            chunk_of<S, TS, T_VX>(a, c, j0, k0, plane, ck[r]);
            chunk_of<S, TS, T_K>(a, c, j0, k0, plane, ck[r]);
        }

        const int i0 = chunk * a.chunk_len;
        const int i1 = min(a.nx2, i0 + a.chunk_len);
        if (i0 >= a.nx2) continue;   // uniform; the host never makes empty chunks
        const int ist = i0 > 0 ? i0 - 1 : 0;
        const int plast = a.nx2 - 1;

        // ---- prologue: planes ist and ist+1 into their slots, plane ist+2
        // in flight; slot of plane p is p mod 3
        int s0 = ist % 3;                     // slot of plane i
        int s1 = s0 == 2 ? 0 : s0 + 1;        // plane i+1
        int s2 = s1 == 2 ? 0 : s1 + 1;        // plane i+2
        u32x4 v[NR];
        stage_issue<NR>(ck, ist, v);
        stage_write<NR>(ck, s_tile + s0 * TS::slot_bytes, v);
        stage_issue<NR>(ck, min(ist + 1, plast), v);
        stage_write<NR>(ck, s_tile + s1 * TS::slot_bytes, v);
        if (ist + 2 <= i1) stage_issue<NR>(ck, min(ist + 2, plast), v);
        Psi mc;
        load_psi<S>(a, L, ist, plane, mc);
        __syncthreads();

        float vx_prev = 0.f;
        for (int i = ist; i < i1; ++i) {
            Psi mn{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            if (i + 1 < i1) load_psi<S>(a, L, i + 1, plane, mn);
            tile_plane<S, TS>(a, L, i, i0, plane, s_tile + s0 * TS::slot_bytes, s_tile + s1 * TS::slot_bytes, mc,
                          hz, vx_prev);
            // plane i+2 into the slot plane i-1 used (read last before the
            // previous barrier); it is read from the next plane on
            if (i + 2 <= i1) stage_write<NR>(ck, s_tile + s2 * TS::slot_bytes, v);
            __syncthreads();
            if (i + 3 <= i1) stage_issue<NR>(ck, min(i + 3, plast), v);
            mc = mn;
            const int t = s0;
            s0 = s1;
            s1 = s2;
            s2 = t;
        }
    }
}

template <typename S, int TJ>
static const void *tile_kernel_ptr()
{
    return reinterpret_cast<const void *>(&a3d_tile_step<S, TJ>);
}

template <typename S>
int tile_max_blocks_per_cu(int tj, int *blocks)
{
    const void *f = nullptr;
    switch (tj) {
    case 1: f = tile_kernel_ptr<S, 1>(); break;
    case 2: f = tile_kernel_ptr<S, 2>(); break;
    case 4: f = tile_kernel_ptr<S, 4>(); break;
    case 8: f = tile_kernel_ptr<S, 8>(); break;
    default: return SM_ERR_ARG;
    }
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64 * tj, 0) != hipSuccess)
        return SM_ERR_HIP;
    *blocks = nb;
    return SM_OK;
}

template <typename S>
int tile_launch(const FusedArgs<S> &a, int tj, hipStream_t st)
{
    const int blocks = a.nblocks + a.gather_blocks;
    switch (tj) {
    case 1: a3d_tile_step<S, 1><<<blocks, dim3(64, 1), 0, st>>>(a); break;
    case 2: a3d_tile_step<S, 2><<<blocks, dim3(64, 2), 0, st>>>(a); break;
    case 4: a3d_tile_step<S, 4><<<blocks, dim3(64, 4), 0, st>>>(a); break;
    case 8: a3d_tile_step<S, 8><<<blocks, dim3(64, 8), 0, st>>>(a); break;
    default: return SM_ERR_ARG;
    }
    return hipGetLastError() == hipSuccess ? SM_OK : SM_ERR_HIP;
}

template int tile_launch<float>(const FusedArgs<float> &, int, hipStream_t);
template int tile_launch<__hip_bfloat16>(const FusedArgs<__hip_bfloat16> &, int, hipStream_t);
template int tile_launch<__half>(const FusedArgs<__half> &, int, hipStream_t);
template int tile_max_blocks_per_cu<float>(int, int *);
template int tile_max_blocks_per_cu<__hip_bfloat16>(int, int *);
template int tile_max_blocks_per_cu<__half>(int, int *);

}  // namespace sm
