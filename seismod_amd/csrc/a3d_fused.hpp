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

template <typename S>
struct FusedArgs {
    int nx2, ny2, nz2, np;
    int pz;               // device row pitch along z (cells, a multiple of 64)
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
    // material: rho at 0, K*dt at nnn
    const float *mat;
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

// pf: planes prefetched ahead (1 or 2); a64: 64-cell aligned z tiles
template <typename S>
int fused_launch(const FusedArgs<S> &a, int tj, int pf, bool a64, hipStream_t st);
template <typename S>
int gather_launch(const FusedArgs<S> &a, hipStream_t st);
// resident blocks per CU of the stencil kernel with TJ = tj
template <typename S>
int fused_max_blocks_per_cu(int tj, int pf, bool a64, int *blocks);
// the same for the two-cells-per-lane kernel (a3d_pair.hip)
template <typename S>
int pair_launch(const FusedArgs<S> &a, int tj, hipStream_t st);
template <typename S>
int pair_max_blocks_per_cu(int tj, int *blocks);

}  // namespace sm
