// a3d_common.hpp -- device-side description of a 3D acoustic CPML solver.
//
// Device layout (all arrays z fastest, like the host model):
//   P, vx, vy, vz, rho, kdt : nx2*ny2*nz2 cells, index (i*ny2+j)*nz2+k
//   CPML memory variables live only in the absorbing slabs:
//     x slabs: [2np][ny2][nz2], y slabs: [nx2][2np][nz2], z slabs: [nx2][ny2][2np]
//   slab s < np is the low side (index n = s), s >= np the high side.
//   The velocity memory variable of axis d (mem_dPd of a_iso_3d.c) covers
//   n < np and n2-1-np <= n < n2-1; the pressure one (mem_dvd) covers
//   1 <= n < np and n2-np <= n < n2.  Outside the slabs the reference's
//   tables are a=0, b=1, kappa=1, which leaves its memory variables at 0 and
//   its update identical to the plain one, so the plain update is used.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "hip/hip_bf16.h"
#include "hip/hip_fp16.h"

namespace sm {

// CPML tables of one axis on the device.  For the FP64 path k holds kappa
// (the reference divides by it); for the FP32 paths it holds 1/kappa.
template <typename R>
struct CpmlAxis {
    const R *a, *b, *k;      // node tables, n2 entries (pressure update)
    const R *ah, *bh, *kh;   // half-node tables, n2-1 entries (velocity update)
};

template <typename T>  // T = storage type
struct A3dDev {
    int nx2, ny2, nz2, np;
    T *P, *vx, *vy, *vz;            // wavefields
    const T *rho, *kdt;             // density and K*dt (FP32 paths: FP32 copies)
    T *mPx, *mPy, *mPz;             // velocity memory variables (slabs)
    T *mvx, *mvy, *mvz;             // pressure memory variables (slabs)
};

// slab coordinate of index n on an axis of n2 nodes, or -1 outside
__host__ __device__ inline int slab_vel(int n, int n2, int np)
{
    if (n < np) return n;
    if (n >= n2 - 1 - np && n < n2 - 1) return np + (n - (n2 - 1 - np));
    return -1;
}
__host__ __device__ inline int slab_pre(int n, int n2, int np)
{
    if (n < np) return n;
    if (n >= n2 - np) return np + (n - (n2 - np));
    return -1;
}

}  // namespace sm
