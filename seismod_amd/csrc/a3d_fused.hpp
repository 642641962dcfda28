// a3d_fused.hpp -- argument block of the fused 3D step kernel.
//
// Kept small on purpose: every field is uniform and lives in SGPRs for the
// whole kernel, so arrays are passed as a few base pointers plus offsets.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "hip/hip_bf16.h"
#include "hip/hip_fp16.h"

namespace sm {

// per-plane constants of the x axis, read with one scalar load per plane
struct XPlane {
    float bh, ah, ikh;   // half-node CPML coefficients (velocity at i)
    float b, a, ik;      // node CPML coefficients (pressure at i)
    int sxv, sxp;        // x slab coordinates (velocity, pressure), -1 outside
    int q0, q1;          // sources of this plane: sorted entries [q0, q1)
    int pad[6];
};

// Plane record i through the constant address space: the record is
// read-only for the whole launch and i is wave-uniform (readfirstlane makes
// that visible to the compiler), so this compiles to one scalar load
// (s_load_dwordx16) instead of vector loads and a wait on them.
__device__ __forceinline__ XPlane xplane(const XPlane *p, int i)
{
    i = __builtin_amdgcn_readfirstlane(i);
    const __attribute__((address_space(4))) int *w =
        (const __attribute__((address_space(4))) int *)(p + i);
    XPlane x;
    x.bh = __int_as_float(w[0]); x.ah = __int_as_float(w[1]); x.ikh = __int_as_float(w[2]);
    x.b = __int_as_float(w[3]); x.a = __int_as_float(w[4]); x.ik = __int_as_float(w[5]);
    x.sxv = w[6]; x.sxp = w[7]; x.q0 = w[8]; x.q1 = w[9];
    return x;
}

// field q (0 P, 1 vx, 2 vy, 3 vz) of cell o: element q*nnn + o of in/out;
// material: rho at o, K*dt at nnn + o
#define F_VX 1u
#define F_VY 2u
#define F_VZ 3u

// z-halo constants of a 64-cell tile (uniform): CPML of the velocity at k0-1
struct Halo {
    float bh, ah, ikh;
};

template <typename S>
struct FusedArgs {
    int nx2, ny2, nz2, np;
    int pz;               // device row pitch along z (cells, >= nz2)
    int ntk, ntiles, chunk_len;
    int nreal;            // tiles * chunks
    int nitems;           // nreal rounded up to a multiple of 8
    int nblocks;          // persistent stencil blocks (grid = nblocks + gather_blocks)
    int gather_blocks;
    unsigned nnn;         // cells per field on the device (nx2 * ny2 * pz)
    // state: P, vx, vy, vz at element q*nnn of in / out (ping-pong)
    const S *in;
    S *out;
    // memory variables: x slabs at 0, y slabs at my, z slabs at mz
    // (velocity set ping-ponged, pressure set updated in place)
    const S *mem_in;
    S *mem_out;
    S *memp;
    unsigned my, mz;
    // material: rho at 0, K*dt at nnn; FP32, or FP16 (tile kernels with
    // 16-bit material) whose value times sR / sK is the physical one
    const float *mat;
    float sR, sK;
    // per-plane x constants, nx2 records
    const XPlane *xpl;
    // y and z CPML tables at tab + toff: a, b, 1/kappa, ah, bh, 1/kappa_h
    // (n2 entries each)
    const float *tab;
    int toff_y, toff_z;
    float idx, idy, idz, two_dt;
    // per-field power-of-two scales: stored value * s = physical value
    // (P, velocities, velocity-side and pressure-side memory variables)
    float sP, isP, sV, isV, sMv, isMv, sMp, isMp;
    // sources sorted by cell: cell, off, len (int64, nsrc each) at src,
    // increments (float) at src_inc
    const int64_t *src;
    const float *src_inc;
    int nsrc;
    int64_t step;
    unsigned salt;   // per-step salt of the stochastic rounding of 16-bit stores
    int sr;          // 1: stochastic rounding of 16-bit stores
    // receivers of the previous step (used by the gather blocks only)
    int nrec;
    const int64_t *ridx;
    const int *rcomp;
    float *trace;
    int64_t gather_row;   // -1: nothing to gather
};

// pf: planes prefetched ahead (1 or 2)
template <typename S>
int fused_launch(const FusedArgs<S> &a, int tj, int pf, hipStream_t st);
template <typename S>
int gather_launch(const FusedArgs<S> &a, hipStream_t st);
// resident blocks per CU of the stencil kernel with TJ = tj
template <typename S>
int fused_max_blocks_per_cu(int tj, int pf, int *blocks);
// the same for the LDS-staged tile kernel (a3d_tile.hip; split: slab
// groups decided per tile and range of planes; nt: state stores
// non-temporal, built for tj = 4 and 8; m16: FP16 material, built for
// 16-bit state, tj = 4 and 8, round-to-nearest stores); the device row
// pitch must be a multiple of 8 cells
template <typename S>
int tile_launch(const FusedArgs<S> &a, int tj, bool split, bool nt, bool m16, hipStream_t st);
template <typename S>
int tile_max_blocks_per_cu(int tj, bool split, bool nt, bool sr, bool m16, int *blocks);
// the FP16-material kernels (a3d_tile_m16.hip)
template <typename S>
int tile_launch_m16(const FusedArgs<S> &a, int tj, bool split, bool nt, hipStream_t st);
template <typename S>
int tile_max_blocks_per_cu_m16(int tj, bool split, bool nt, bool sr, int *blocks);

}  // namespace sm
