// a3d_fused.hpp -- argument block of the fused 3D step kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "hip/hip_bf16.h"
#include "hip/hip_fp16.h"

namespace sm {

template <typename S>
struct FusedArgs {
    int nx2, ny2, nz2, np;
    int ntk, ntiles, chunk_len;
    int nreal;            // tiles * chunks
    int nitems;           // nreal rounded up to a multiple of 8
    int nblocks;          // persistent stencil blocks (grid = nblocks + gather_blocks)
    int gather_blocks;
    // state in/out (ping-pong) and memory variables
    const S *P_in, *vx_in, *vy_in, *vz_in;
    S *P_out, *vx_out, *vy_out, *vz_out;
    const S *mPx_in, *mPy_in, *mPz_in;
    S *mPx_out, *mPy_out, *mPz_out;
    S *mvx, *mvy, *mvz;                    // pressure memory variables, in place
    const float *rho, *kdt;
    // CPML tables (node: a, b, 1/kappa; half node: ah, bh, 1/kappa_h)
    const float *ax, *bx, *ikx, *ay, *by, *iky, *az, *bz, *ikz;
    const float *ahx, *bhx, *ikhx, *ahy, *bhy, *ikhy, *ahz, *bhz, *ikhz;
    float idx, idy, idz, two_dt;
    // per-field power-of-two scales: stored value * s = physical value
    float sP, isP, sV, isV, sM, isM;
    // sources sorted by cell; src_plane[i]..src_plane[i+1] lists plane i
    const int *src_plane;
    const int64_t *src_cell, *src_off, *src_len;
    const float *src_inc;
    int64_t step;
    // receivers of the previous step
    int nrec;
    const int64_t *ridx;
    const int *rcomp;
    float *trace;
    int64_t gather_row;   // -1: nothing to gather
};

template <typename S>
int fused_launch(const FusedArgs<S> &a, int tj, hipStream_t st);
template <typename S>
int gather_launch(const FusedArgs<S> &a, hipStream_t st);
// resident blocks per CU of the stencil kernel with TJ = tj
template <typename S>
int fused_max_blocks_per_cu(int tj, int *blocks);

}  // namespace sm
