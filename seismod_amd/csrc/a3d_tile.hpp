// a3d_tile.hpp -- the LDS-staged 3D acoustic step kernel (a3d_tile_step)
// and its helpers; a3d_tile.hip describes the design and launches the
// kernels with FP32 material, a3d_tile_m16.hip those with 16-bit material
// (separate translation units, so the two sets compile in parallel).
#pragma once
#include <cstdint>

#include "a3d_common.hpp"
#include "a3d_fused.hpp"
#include "a3d_store.hpp"
#include "seismod_hip.h"

namespace sm {

namespace {

constexpr unsigned OFF_NONE = 0xFFFFFF00u;   // past the end of every buffer
constexpr int RSRC_WORD3 = 0x00020000;       // raw 32-bit buffer on gfx9

// cache policy (buffer instruction aux bits) of the state stores when
// they are non-temporal: streamed past the caches
constexpr int CPOL_NT = 2;

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
using Rsrc = __amdgpu_buffer_rsrc_t;

__device__ __forceinline__ Rsrc rsrc(const void *p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, RSRC_WORD3);
}

// load / store one element of storage type S at byte offset voff
template <typename S>
__device__ __forceinline__ float bld(Rsrc r, unsigned voff, float scale)
{
    if constexpr (sizeof(S) == 4)
        return Store<S>::dec(__builtin_amdgcn_raw_buffer_load_b32(r, voff, 0, 0), scale);
    else
        return Store<S>::dec(__builtin_amdgcn_raw_buffer_load_b16(r, voff, 0, 0), scale);
}
template <typename S>
__device__ __forceinline__ void bst(Rsrc r, unsigned voff, typename Store<S>::Bits b)
{
    if constexpr (sizeof(S) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(b, r, voff, 0, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b16(b, r, voff, 0, 0);
}


// Tile fields of one plane slot.  The fields of the state buffer come
// first, those of the material buffer start at a multiple of 64 chunks, so
// every wave copies from one buffer per round.
enum { T_P = 0, T_VY = 1, T_VZ = 2, T_VX = 3, T_R = 4, T_K = 5, T_NF = 6 };

// Geometry of the staged tiles for storage type S and TJ rows per block.
// Rows are relative to j0 (the block's first row), cells to k0 (first
// cell).  Chunk c of a slot lives at byte c * 16 of the slot.  SR: 16-bit
// stores round stochastically (a compile-time choice, so the default
// round-to-nearest kernels carry no branch or salt for it).  NTS: the
// state stores are non-temporal.  M16: the material (rho, K*dt) is stored
// as FP16 with a power-of-two scale per field instead of FP32.
template <typename S, int TJ, bool SR = false, bool NTS = false, bool M16 = false>
struct TileShape {
    static constexpr int sr = SR ? 1 : 0;
    static constexpr int st_cpol = NTS ? CPOL_NT : 0;
    static constexpr int ES = (int)sizeof(S);
    static constexpr int MES = M16 ? 2 : 4;                           // material element
    static constexpr int esz(int f) { return f >= T_R ? MES : ES; }
    static constexpr int cpc(int f) { return 16 / esz(f); }   // cells per 16-byte chunk
    static constexpr int rows(int f)
    {
        return (f == T_P || f == T_R) ? TJ + 2 : f == T_VY ? TJ + 1 : TJ;
    }
    static constexpr int row0(int f) { return (f == T_P || f == T_R || f == T_VY) ? -1 : 0; }
    static constexpr int c0(int f) { return (f == T_P || f == T_R || f == T_VZ) ? -cpc(f) : 0; }
    static constexpr int cells(int f)
    {
        return (f == T_P || f == T_R) ? 64 + 2 * cpc(f) : f == T_VZ ? 64 + cpc(f) : 64;
    }
    static constexpr int row_bytes(int f) { return cells(f) * esz(f); }
    static constexpr int cpr(int f) { return row_bytes(f) / 16; }   // chunks per row
    static constexpr int chunks(int f) { return rows(f) * cpr(f); }
    static constexpr int state_chunks = chunks(T_P) + chunks(T_VY) + chunks(T_VZ) + chunks(T_VX);
    static constexpr int MAT0 = (state_chunks + 63) / 64 * 64;
    static constexpr int chunk0(int f)
    {
        return f == T_P ? 0 : f == T_VY ? chunks(T_P) : f == T_VZ ? chunks(T_P) + chunks(T_VY)
             : f == T_VX ? chunks(T_P) + chunks(T_VY) + chunks(T_VZ)
             : f == T_R ? MAT0 : MAT0 + chunks(T_R);
    }
    static constexpr int off(int f) { return chunk0(f) * 16; }
    static constexpr int total_chunks = MAT0 + chunks(T_R) + chunks(T_K);
    static constexpr int NT = 64 * TJ;                                // threads per block
    static constexpr int NR = (total_chunks + NT - 1) / NT;          // chunks per thread
    static constexpr int slot_bytes = NR * NT * 16;                   // every thread's chunks
    // packed stores: a wave's new P, vx, vy, vz (64 cells each) gathered in
    // LDS, then written as 16-byte chunks, CPF chunks per field
    static constexpr int CPF = 64 * ES / 16;
    static constexpr int stash_bytes = 4 * 64 * ES;                  // per wave
    static constexpr int lds_bytes = 3 * slot_bytes + TJ * stash_bytes;
};

// Per-item constants of one lane (the cell (j, kk) it owns, clamped copies,
// slab coordinates, CPML coefficients, tile coordinates of its reads).
struct TLane {
    int j, kk, kc;
    bool own;
    unsigned o_jk;                       // in-plane device offset of the cell
    int syv, syp, syh, szv, szp, szh;    // slab coordinates (-1: not in a slab)
    unsigned mz_v, mz_p, mz_h;           // z-slab offsets within a plane
    float ybh, yah, yikh, yb, ya, yik, hbh, hah, hikh;
    float zbh, zah, zikh, zb, za, zik;
    // rows (relative to j0, wave-uniform) of the cell, j+1 and j-1;
    // columns (relative to k0) of the cell and k+1 (per lane), of k0-1
    int tc, tjp, tjm;
    int cc, ckp, ckm;
    // packed stores: element offset (plane 0) of the lane's 16-byte chunk
    // of the new state, and whether the lane writes one
    unsigned pk_off;
    bool pk_ok;
};

// CPML memory variables one lane reads for one plane (0 outside the slabs)
struct Psi {
    float mPx, mPy, mPz;   // velocity set at the own cell
    float mPym;            // velocity set (y) at row j-1
    float mPzm;            // velocity set (z) at k0-1 (lane 0)
    float mvx, mvy, mvz;   // pressure set at the own cell
};

// the buffers of one launch
struct Bufs {
    Rsrc in, mat, out, mi, mo, mp;
};

// LDS byte offset within a slot of field F at tile row t, column c
template <typename TS, int F>
__device__ __forceinline__ unsigned tof(int t, int c)
{
    constexpr int O = TS::off(F), R0 = TS::row0(F), RB = TS::row_bytes(F), C0 = TS::c0(F),
                  E = TS::esz(F);
    return (unsigned)(O + (t - R0) * RB + (c - C0) * E);
}

// LDS of one wave written by its lanes and then read by other lanes of the
// same wave: LDS operations of a wave complete in order, so only the
// compiler must not move them across this point
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename S>
__device__ __forceinline__ void stash_put(char *stash, int q, typename Store<S>::Bits b)
{
    ((typename Store<S>::Bits *)stash)[q * 64 + threadIdx.x] = b;
}

template <typename S>
__device__ __forceinline__ float lds_state(const char *slot, unsigned off, float scale)
{
    return Store<S>::ld((const S *)(slot + off), 0u, scale);
}
__device__ __forceinline__ float lds_f32(const char *slot, unsigned off)
{
    return *(const float *)(slot + off);
}
// material value (rho or K*dt) from a slot: FP32, or FP16 times its scale
template <typename TS>
__device__ __forceinline__ float lds_mat(const char *slot, unsigned off, float scale)
{
    if constexpr (TS::MES == 4) return lds_f32(slot, off);
    else return __half2float(*(const __half *)(slot + off)) * scale;
}

// byte offset (within its buffer, plane 0) of chunk c if it belongs to
// field F.  Tile rows are clamped to the grid; chunks past the row end (no
// cell of the grid) and before k = 0 (halo of the first tile, never read)
// are clamped into the row.
template <typename S, typename TS, int F>
__device__ __forceinline__ void chunk_of(const FusedArgs<S> &a, int c, int j0, int k0, unsigned &off)
{
    constexpr int C0 = TS::chunk0(F), NC = TS::chunks(F), CPR = TS::cpr(F), CPC = TS::cpc(F),
                  R0 = TS::row0(F), Z0 = TS::c0(F), E = TS::esz(F);
    if (c < C0 || c >= C0 + NC) return;
    const int q = c - C0;
    const int row = q / CPR, cc = q - row * CPR;
    const int jr = min(max(j0 + R0 + row, 0), a.ny2 - 1);
    int cs = k0 + Z0 + cc * CPC;
    cs = cs < 0 ? 0 : cs >= a.pz ? a.pz - CPC : cs;
    unsigned fo = 0;   // field offset within its buffer, elements
    if constexpr (F == T_K) fo = a.nnn;
    else if constexpr (F == T_VX) fo = F_VX * a.nnn;
    else if constexpr (F == T_VY) fo = F_VY * a.nnn;
    else if constexpr (F == T_VZ) fo = F_VZ * a.nnn;
    off = (fo + (unsigned)jr * a.pz + (unsigned)cs) * (unsigned)E;
}

// the chunk loads of plane p into v (OFF_NONE when p is not needed)
template <typename TS, int NR>
__device__ __forceinline__ void stage_issue(const Bufs &B, const unsigned (&coff)[NR], int w,
                                            unsigned plane, int p, int pmax, int plast,
                                            u32x4 (&v)[NR])
{
    const unsigned pc = (unsigned)min(p, plast);
    const bool need = p <= pmax;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const bool mat = r * TS::NT + w * 64 >= TS::MAT0;   // wave-uniform
        const unsigned ps = plane * (unsigned)(mat ? TS::MES : TS::ES);
        const unsigned vo = need && coff[r] != OFF_NONE ? coff[r] + pc * ps : OFF_NONE;
        v[r] = __builtin_amdgcn_raw_buffer_load_b128(mat ? B.mat : B.in, vo, 0, 0);
    }
}

template <typename TS, int NR>
__device__ __forceinline__ void stage_write(char *slot, int tid, const u32x4 (&v)[NR])
{
#pragma unroll
    for (int r = 0; r < NR; ++r) *(u32x4 *)(slot + (r * TS::NT + tid) * 16) = v[r];
}

// Memory variables of plane p.  X, Y, Z: whether the x, y, z groups can
// be non-zero here (set outside the plane loops, so every loop issues a
// fixed set of loads); a group that is off is 0 without any load.
template <typename S, bool X, bool Y, bool Z>
__device__ __forceinline__ void psi_issue(const FusedArgs<S> &a, const Bufs &B, const TLane &L, int p,
                                          bool need, unsigned plane, Psi &m)
{
    const XPlane x = xplane(a.xpl, min(p, a.nx2 - 1));
    const unsigned np2 = 2u * a.np;
    const unsigned yrow = (unsigned)p * np2;
    const unsigned zpl = a.mz + (unsigned)p * a.ny2 * np2;
    constexpr unsigned E = sizeof(S);
    auto o = [&](bool ok, unsigned idx) { return need && ok ? idx * E : OFF_NONE; };
    m.mPx = m.mvx = m.mPy = m.mvy = m.mPym = m.mPz = m.mvz = m.mPzm = 0.f;
    if constexpr (X) {
        m.mPx = bld<S>(B.mi, o(x.sxv >= 0, (unsigned)x.sxv * plane + L.o_jk), a.sMv);
        m.mvx = bld<S>(B.mp, o(x.sxp >= 0, (unsigned)x.sxp * plane + L.o_jk), a.sMp);
    }
    if constexpr (Y) {
        m.mPy = bld<S>(B.mi, o(L.syv >= 0, a.my + (yrow + L.syv) * a.pz + L.kc), a.sMv);
        m.mPym = bld<S>(B.mi, o(L.syh >= 0, a.my + (yrow + L.syh) * a.pz + L.kc), a.sMv);
        m.mvy = bld<S>(B.mp, o(L.syp >= 0, a.my + (yrow + L.syp) * a.pz + L.kc), a.sMp);
    }
    if constexpr (Z) {
        m.mPz = bld<S>(B.mi, o(L.szv >= 0, zpl + L.mz_v), a.sMv);
        m.mPzm = bld<S>(B.mi, o(L.szh >= 0, zpl + L.mz_h), a.sMv);
        m.mvz = bld<S>(B.mp, o(L.szp >= 0, zpl + L.mz_p), a.sMp);
    }
}

// vx of plane i only (the plane before a chunk's first plane)
template <typename S, typename TS>
__device__ __forceinline__ float tile_vx(const FusedArgs<S> &a, const TLane &L, int i,
                                         const char *cs, const char *ns, const Psi &m)
{
    const XPlane x = xplane(a.xpl, i);
    const unsigned p_c = tof<TS, T_P>(L.tc, L.cc), r_c_ = tof<TS, T_R>(L.tc, L.cc);
    float mx = m.mPx;
    return vel_update(lds_state<S>(cs, tof<TS, T_VX>(L.tc, L.cc), a.sV), lds_state<S>(cs, p_c, a.sP),
                      lds_state<S>(ns, p_c, a.sP), lds_mat<TS>(cs, r_c_, a.sR), lds_mat<TS>(ns, r_c_, a.sR), a.idx, a.two_dt,
                      x.sxv >= 0, x.bh, x.ah, x.ikh, mx);
}

// update plane i from slot cs (plane i) and ns (plane i+1); same expressions
// and operand order as the other variants' plane update.  X, Y, Z: whether
// plane i / the tile can have x, y, z slab cells (when off, every slab test
// of that direction is false here, so its code is left out).
template <typename S, typename TS, bool X, bool Y, bool Z>
__device__ __forceinline__ void tile_plane(const FusedArgs<S> &a, const Bufs &B, const TLane &L, int i,
                                           unsigned plane, const char *cs, const char *ns,
                                           const Psi &m, const Halo &hz, float &vx_prev,
                                           char *stash)
{
    constexpr unsigned E = sizeof(S);
    const XPlane x = xplane(a.xpl, i);
    const unsigned o = (unsigned)i * plane + L.o_jk;
    const unsigned np2 = 2u * a.np;
    const unsigned p_c = tof<TS, T_P>(L.tc, L.cc), r_c_ = tof<TS, T_R>(L.tc, L.cc);
    const float P_c = lds_state<S>(cs, p_c, a.sP);
    const float r_c = lds_mat<TS>(cs, r_c_, a.sR);

    // vx at (i, j, k), a_iso_3d.c:605-619 (the next plane clamped to the last)
    float mx = m.mPx;
    const float vxn = vel_update(lds_state<S>(cs, tof<TS, T_VX>(L.tc, L.cc), a.sV), P_c,
                                 lds_state<S>(ns, p_c, a.sP), r_c, lds_mat<TS>(ns, r_c_, a.sR), a.idx,
                                 a.two_dt, X && x.sxv >= 0, x.bh, x.ah, x.ikh, mx);
    // vy, vz at (i, j, k), a_iso_3d.c:621-651
    float my = m.mPy, mz = m.mPz;
    const float vyn = vel_update(lds_state<S>(cs, tof<TS, T_VY>(L.tc, L.cc), a.sV), P_c,
                                 lds_state<S>(cs, tof<TS, T_P>(L.tjp, L.cc), a.sP), r_c,
                                 lds_mat<TS>(cs, tof<TS, T_R>(L.tjp, L.cc), a.sR), a.idy, a.two_dt, Y && L.syv >= 0,
                                 L.ybh, L.yah, L.yikh, my);
    const float vzn = vel_update(lds_state<S>(cs, tof<TS, T_VZ>(L.tc, L.cc), a.sV), P_c,
                                 lds_state<S>(cs, tof<TS, T_P>(L.tc, L.ckp), a.sP), r_c,
                                 lds_mat<TS>(cs, tof<TS, T_R>(L.tc, L.ckp), a.sR), a.idz, a.two_dt, Z && L.szv >= 0,
                                 L.zbh, L.zah, L.zikh, mz);
    // the new velocities go to the wave's LDS stash (written out with P
    // below as 16-byte chunks)
    stash_put<S>(stash, 1, Store<S>::enc(vxn, a.isV, F_VX * a.nnn + o, a.salt, TS::sr));
    stash_put<S>(stash, 2, Store<S>::enc(vyn, a.isV, F_VY * a.nnn + o, a.salt, TS::sr));
    stash_put<S>(stash, 3, Store<S>::enc(vzn, a.isV, F_VZ * a.nnn + o, a.salt, TS::sr));
    // memory-variable stores: a lane that must not write passes OFF_NONE
    {
        const unsigned ix = (unsigned)x.sxv * plane + L.o_jk;
        const unsigned iy = a.my + ((unsigned)i * np2 + L.syv) * a.pz + L.kc;
        const unsigned iz = a.mz + (unsigned)i * a.ny2 * np2 + L.mz_v;
        if constexpr (X)
            bst<S>(B.mo, L.own && x.sxv >= 0 ? ix * E : OFF_NONE,
                   Store<S>::enc(mx, a.isMv, ix, a.salt ^ SALT_MV, TS::sr));
        if constexpr (Y)
            bst<S>(B.mo, L.own && L.syv >= 0 ? iy * E : OFF_NONE,
                   Store<S>::enc(my, a.isMv, iy, a.salt ^ SALT_MV, TS::sr));
        if constexpr (Z)
            bst<S>(B.mo, L.own && L.szv >= 0 ? iz * E : OFF_NONE,
                   Store<S>::enc(mz, a.isMv, iz, a.salt ^ SALT_MV, TS::sr));
    }
    // new vy of row j-1 and new vz of k0-1 (owned by other waves / tiles;
    // recomputed here, not stored)
    float mh = m.mPym;
    const float vy_m = vel_update(lds_state<S>(cs, tof<TS, T_VY>(L.tjm, L.cc), a.sV),
                                  lds_state<S>(cs, tof<TS, T_P>(L.tjm, L.cc), a.sP), P_c,
                                  lds_mat<TS>(cs, tof<TS, T_R>(L.tjm, L.cc), a.sR), r_c, a.idy, a.two_dt,
                                  Y && L.syh >= 0, L.hbh, L.hah, L.hikh, mh);
    float vz_m = __shfl_up(vzn, 1);
    {
        float mzh = m.mPzm;
        const float vz_h = vel_update(lds_state<S>(cs, tof<TS, T_VZ>(L.tc, L.ckm), a.sV),
                                      lds_state<S>(cs, tof<TS, T_P>(L.tc, L.ckm), a.sP), P_c,
                                      lds_mat<TS>(cs, tof<TS, T_R>(L.tc, L.ckm), a.sR), r_c, a.idz, a.two_dt,
                                      Z && L.szh >= 0, hz.bh, hz.ah, hz.ikh, mzh);
        if (threadIdx.x == 0) vz_m = vz_h;
    }

    // pressure at (i, j, k), a_iso_3d.c:659-683; memory variables of the
    // pressure set are written only where the reference updates them
    const bool upd = L.own && i >= 1 && L.j >= 1 && L.kk >= 1;
    float dvdx = (vxn - vx_prev) * a.idx;
    float dvdy = (vyn - vy_m) * a.idy;
    float dvdz = (vzn - vz_m) * a.idz;
    const unsigned jx = (unsigned)x.sxp * plane + L.o_jk;
    const unsigned jy = a.my + ((unsigned)i * np2 + L.syp) * a.pz + L.kc;
    const unsigned jz = a.mz + (unsigned)i * a.ny2 * np2 + L.mz_p;
    float mmx = 0.f, mmy = 0.f, mmz = 0.f;
    if (X && x.sxp >= 0) {
        mmx = x.b * m.mvx + x.a * dvdx;
        dvdx = dvdx * x.ik + mmx;
    }
    if (Y && L.syp >= 0) {
        mmy = L.yb * m.mvy + L.ya * dvdy;
        dvdy = dvdy * L.yik + mmy;
    }
    if (Z && L.szp >= 0) {
        mmz = L.zb * m.mvz + L.za * dvdz;
        dvdz = dvdz * L.zik + mmz;
    }
    if constexpr (X)
        bst<S>(B.mp, upd && x.sxp >= 0 ? jx * E : OFF_NONE,
               Store<S>::enc(mmx, a.isMp, jx, a.salt ^ SALT_MP, TS::sr));
    if constexpr (Y)
        bst<S>(B.mp, upd && L.syp >= 0 ? jy * E : OFF_NONE,
               Store<S>::enc(mmy, a.isMp, jy, a.salt ^ SALT_MP, TS::sr));
    if constexpr (Z)
        bst<S>(B.mp, upd && L.szp >= 0 ? jz * E : OFF_NONE,
               Store<S>::enc(mmz, a.isMp, jz, a.salt ^ SALT_MP, TS::sr));
    float Pn = upd ? P_c - lds_mat<TS>(cs, tof<TS, T_K>(L.tc, L.cc), a.sK) * (dvdx + dvdy + dvdz) : P_c;
    if (x.q0 < x.q1) {
        // sources of this plane (a_iso_3d.c:689-693), in source order, from
        // the source table through scalar loads
        typedef const __attribute__((address_space(4))) int64_t *C64;
        typedef const __attribute__((address_space(4))) float *CF;
        for (int q = x.q0; q < x.q1; ++q) {
            const int64_t cell = ((C64)a.src)[q];
            const int64_t off = ((C64)a.src)[a.nsrc + q];
            const int64_t len = ((C64)a.src)[2 * a.nsrc + q];
            if (L.own && (unsigned)cell == o && a.step < len) Pn += ((CF)a.src_inc)[off + a.step];
        }
    }
    // the wave's four new rows leave as 16-byte stores, lane l taking
    // chunk l % CPF of field l / CPF (lanes past the four fields and rows
    // past the grid pass OFF_NONE); cells past the grid edge in a row (pad
    // cells) receive the values their lanes computed, which no update reads
    stash_put<S>(stash, 0, Store<S>::enc(Pn, a.isP, o, a.salt, TS::sr));
    wave_lds_sync();
    const u32x4 pk = *(const u32x4 *)(stash + (threadIdx.x % (4 * TS::CPF)) * 16);
    __builtin_amdgcn_raw_buffer_store_b128(pk, B.out, L.pk_ok ? (L.pk_off + i * plane) * E : OFF_NONE,
                                           0, TS::st_cpol);
    wave_lds_sync();
    vx_prev = vxn;
}

// Loop state of a chunk [i0, i1).  At plane i the register set of the
// parity of i - i0 holds the chunks of plane i+2, and the memory-variable
// set of that parity (m0 even, m1 odd) those of plane i while the other
// receives plane i+1's; slots s0, s1, s2 hold planes i, i+1 and the free
// one.
struct Ring {
    int s0, s1, s2;
    Psi m0, m1;
    float vx_prev;
};

// one plane of a range: memory variables of plane i+1 issued into mn,
// plane i updated with those in mc, plane i+2 (in v) written to the free
// slot, loads of plane i+4 issued into v.  mc and mn are fixed register
// sets used in turn by the two planes of a trip: a rotation by copy
// (m0 = m1; m1 = m2) made every plane wait for the loads it had just
// issued and, as loads complete in order, for the chunk loads before them.
template <typename S, typename TS, bool X, bool Y, bool Z>
__device__ __forceinline__ void plane_iter(const FusedArgs<S> &a, const Bufs &B, const TLane &L,
                                           const Halo &hz, const unsigned (&coff)[TS::NR], int w,
                                           int tid, unsigned plane, int i1, int i, char *base,
                                           u32x4 (&v)[TS::NR], Ring &R, const Psi &mc, Psi &mn)
{
    constexpr int NR = TS::NR;
    psi_issue<S, X, Y, Z>(a, B, L, i + 1, i + 1 < i1, plane, mn);
    tile_plane<S, TS, X, Y, Z>(a, B, L, i, plane, base + R.s0 * TS::slot_bytes,
                               base + R.s1 * TS::slot_bytes, mc, hz, R.vx_prev,
                               base + 3 * TS::slot_bytes + w * TS::stash_bytes);
    // plane i+2 into the slot plane i-1 used (read last before the previous
    // barrier), then the loads of plane i+4 into the same register set
    stage_write<TS, NR>(base + R.s2 * TS::slot_bytes, tid, v);
    __syncthreads();
    stage_issue<TS, NR>(B, coff, w, plane, i + 4, i1, a.nx2 - 1, v);
    const int t = R.s0;
    R.s0 = R.s1;
    R.s1 = R.s2;
    R.s2 = t;
}

// Planes [ib, ie) of a chunk ending at i1 with the x flag fixed, two per
// trip: v0 holds the chunks of plane ib+2, v1 those of plane ib+3; R.m0
// the memory variables of plane ib (ib - i0 is even).
template <typename S, typename TS, bool X, bool Y, bool Z>
__device__ __forceinline__ void plane_range(const FusedArgs<S> &a, const Bufs &B, const TLane &L,
                                            const Halo &hz, const unsigned (&coff)[TS::NR], int w,
                                            int tid, unsigned plane, int i1, int ib, int ie,
                                            char *base, u32x4 (&v0)[TS::NR], u32x4 (&v1)[TS::NR],
                                            Ring &R)
{
    for (int i = ib; i < ie; i += 2) {
        plane_iter<S, TS, X, Y, Z>(a, B, L, hz, coff, w, tid, plane, i1, i, base, v0, R, R.m0, R.m1);
        if (i + 1 >= ie) break;
        plane_iter<S, TS, X, Y, Z>(a, B, L, hz, coff, w, tid, plane, i1, i + 1, base, v1, R, R.m1,
                                   R.m0);
    }
}

// One work item: planes [i0, i1) of one tile.  Y, Z: the tile has cells
// (or neighbours it recomputes) in a y / z slab.  SPLIT: the chunk is
// marched in up to three ranges, x-slab planes at either end (and the planes
// before the upper ones, which load their memory variables) with the x group
// and the interior without; otherwise in one range with every group on.
template <typename S, typename TS, bool Y, bool Z, bool SPLIT>
__device__ __forceinline__ void march(const FusedArgs<S> &a, const Bufs &B, const TLane &L,
                                      const Halo &hz, const unsigned (&coff)[TS::NR], int w, int tid,
                                      unsigned plane, int i0, int i1, char *base)
{
    constexpr int NR = TS::NR;
    const int pmax = i1, plast = a.nx2 - 1;
    // prologue (synchronous): planes i0, i0+1 and, for a chunk after the
    // first, i0-1 into their slots (plane p in slot p mod 3); vx of plane
    // i0-1
    Ring R;
    R.s0 = i0 % 3;
    R.s1 = R.s0 == 2 ? 0 : R.s0 + 1;
    R.s2 = R.s1 == 2 ? 0 : R.s1 + 1;
    R.vx_prev = 0.f;
    u32x4 va[NR], vb[NR];
    stage_issue<TS, NR>(B, coff, w, plane, i0, pmax, plast, va);
    stage_issue<TS, NR>(B, coff, w, plane, i0 + 1, pmax, plast, vb);
    stage_write<TS, NR>(base + R.s0 * TS::slot_bytes, tid, va);
    stage_write<TS, NR>(base + R.s1 * TS::slot_bytes, tid, vb);
    if (i0 > 0) {
        stage_issue<TS, NR>(B, coff, w, plane, i0 - 1, pmax, plast, va);
        psi_issue<S, true, Y, Z>(a, B, L, i0 - 1, true, plane, R.m0);
        stage_write<TS, NR>(base + R.s2 * TS::slot_bytes, tid, va);
        __syncthreads();
        R.vx_prev = tile_vx<S, TS>(a, L, i0 - 1, base + R.s2 * TS::slot_bytes,
                                   base + R.s0 * TS::slot_bytes, R.m0);
    }
    __syncthreads();   // slots of planes i0, i0+1 written; slot s2 free
    stage_issue<TS, NR>(B, coff, w, plane, i0 + 2, pmax, plast, va);
    psi_issue<S, true, Y, Z>(a, B, L, i0, true, plane, R.m0);
    stage_issue<TS, NR>(B, coff, w, plane, i0 + 3, pmax, plast, vb);
    // x-slab planes: i < np or i >= nx2-1-np (velocity or pressure set);
    // the x group must be on where plane i or plane i+1 (whose memory
    // variables plane i loads) is one, and is on two planes ahead of the
    // upper slab.  The interior range [b1, b2) starts and has a length even
    // in planes from i0, so every range starts with register set va and
    // memory-variable set m0 (the x ranges grow by at most one plane; the x
    // group is correct anywhere)
    const int b1 = min(i0 + (max(a.np - i0, 0) + 1) / 2 * 2, i1);
    const int b2 = b1 + max(min(a.nx2 - 3 - a.np, i1) - b1, 0) / 2 * 2;
    if constexpr (SPLIT) {
        plane_range<S, TS, true, Y, Z>(a, B, L, hz, coff, w, tid, plane, i1, i0, b1, base, va, vb, R);
        plane_range<S, TS, false, Y, Z>(a, B, L, hz, coff, w, tid, plane, i1, b1, b2, base, va, vb, R);
        plane_range<S, TS, true, Y, Z>(a, B, L, hz, coff, w, tid, plane, i1, b2, i1, base, va, vb, R);
    } else {
        plane_range<S, TS, true, Y, Z>(a, B, L, hz, coff, w, tid, plane, i1, i0, i1, base, va, vb, R);
    }
}

}  // namespace

// SPLIT: slab groups decided per tile and per range of planes (more code,
// fewer instructions per interior cell); otherwise one loop with every
// group.  Both are held to at least 4 waves per SIMD (at most 128
// registers per lane).
template <typename S, int TJ, bool SPLIT, bool SR, bool NT, bool M16>
__global__ void __launch_bounds__(64 * TJ, 4)
a3d_tile_step(FusedArgs<S> a)
{
    using TS = TileShape<S, TJ, SR, NT, M16>;
    constexpr int NR = TS::NR;
    __shared__ __attribute__((aligned(16))) char s_tile[TS::lds_bytes];

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
    const unsigned state_bytes = 4u * a.nnn * (unsigned)sizeof(S);
    const unsigned slab_elems = a.mz + (unsigned)a.nx2 * a.ny2 * np2;
    Bufs B;
    B.in = rsrc(a.in, state_bytes);
    B.out = rsrc(a.out, state_bytes);
    B.mat = rsrc(a.mat, 2u * a.nnn * (unsigned)TS::MES);
    B.mi = rsrc(a.mem_in, slab_elems * (unsigned)sizeof(S));
    B.mo = rsrc(a.mem_out, slab_elems * (unsigned)sizeof(S));
    B.mp = rsrc(a.memp, slab_elems * (unsigned)sizeof(S));

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
        L.tc = jc - j0; L.tjp = jp - j0; L.tjm = jm - j0;
        L.cc = L.kc - k0; L.ckp = kp - k0; L.ckm = km - k0;
        {
            // packed stores: lane l writes field l / CPF, cells from
            // (l % CPF) * (16 / ES) of the tile (the pitch is a multiple of
            // 8 cells, so a chunk starting inside the row ends inside it)
            const int q = lane / TS::CPF, c = (lane % TS::CPF) * (16 / TS::ES);
            L.pk_ok = jv && q < 4 && k0 + c < a.pz;
            L.pk_off = (unsigned)q * a.nnn + (unsigned)jc * a.pz + (unsigned)(k0 + c);
        }
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
        // ---- the 16-byte chunks this thread copies (offsets in plane 0)
        unsigned coff[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int c = r * TS::NT + tid;
            coff[r] = OFF_NONE;
            chunk_of<S, TS, T_P>(a, c, j0, k0, coff[r]);
            chunk_of<S, TS, T_VY>(a, c, j0, k0, coff[r]);
            chunk_of<S, TS, T_VZ>(a, c, j0, k0, coff[r]);
            chunk_of<S, TS, T_VX>(a, c, j0, k0, coff[r]);
            chunk_of<S, TS, T_R>(a, c, j0, k0, coff[r]);
            chunk_of<S, TS, T_K>(a, c, j0, k0, coff[r]);
        }

        const int i0 = chunk * a.chunk_len;
        const int i1 = min(a.nx2, i0 + a.chunk_len);
        if (i0 >= a.nx2) continue;   // uniform; the host never makes empty chunks
        // the tile's rows (and the row below) or cells (and k0-1) in a y / z
        // slab: rows j <= np or j >= ny2-1-np, cells likewise
        if constexpr (SPLIT) {
            const bool ys = j0 <= a.np || j0 + TJ - 1 >= a.ny2 - 1 - a.np;
            const bool zs = k0 <= a.np || k0 + 63 >= a.nz2 - 1 - a.np;
            if (ys) {
                if (zs) march<S, TS, true, true, true>(a, B, L, hz, coff, w, tid, plane, i0, i1, s_tile);
                else march<S, TS, true, false, true>(a, B, L, hz, coff, w, tid, plane, i0, i1, s_tile);
            } else {
                if (zs) march<S, TS, false, true, true>(a, B, L, hz, coff, w, tid, plane, i0, i1, s_tile);
                else march<S, TS, false, false, true>(a, B, L, hz, coff, w, tid, plane, i0, i1, s_tile);
            }
        } else {
            march<S, TS, true, true, false>(a, B, L, hz, coff, w, tid, plane, i0, i1, s_tile);
        }
    }
}

}  // namespace sm
