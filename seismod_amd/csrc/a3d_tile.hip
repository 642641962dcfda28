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
//     item order, chunks with i0 > 0 first compute vx of plane i0-1).
//   - for every plane the block copies what it reads into LDS: P and rho
//     with one row and one 16-byte chunk of halo on each side, old vy with
//     the row below, old vz with the chunk to the left, old vx and K*dt.
//     Each lane copies 16-byte chunks (buffer_load_dwordx4, ds_write_b128),
//     so every row of every field crosses the memory pipeline once per
//     plane, in whole chunks, and all neighbour values come from LDS.
//   - three plane slots rotate: plane i is updated from slots i and i+1
//     while plane i+2 is written to the third.  Chunk loads are issued two
//     planes ahead of their use and CPML memory-variable loads one plane
//     ahead, each into two register sets that the two planes of a loop trip
//     use in turn (no value still in flight is ever copied), so HBM latency
//     hides behind two planes of work.  One barrier per plane.
//   - every vector-memory access inside the plane loop is unconditional: a
//     lane with nothing to load or store passes an offset past the end of
//     its buffer resource (OFF_NONE), for which the hardware returns 0 and
//     drops stores without touching memory.  The compiler then knows the
//     exact number of accesses in flight and waits only for the ones a
//     value needs (vmcnt), instead of draining the queue every plane.
//   - each wave recomputes new vy of the row below its own and lane 0 new
//     vz of the cell left of the tile, with the expressions their owners use,
//     so no values pass between waves within a plane.
//
// Requirements: the device row pitch is a multiple of 8 cells (16 bytes of
// 16-bit storage), so chunks never straddle a row end; every buffer is
// smaller than OFF_NONE bytes (checked at creation).
//
// The kernel and its helpers are in a3d_tile.hpp; this file launches the
// kernels that read FP32 material and hands 16-bit material to
// a3d_tile_m16.hip.
#include "a3d_tile.hpp"

namespace sm {

// stochastic rounding exists only for the 16-bit storage types
template <typename S>
constexpr bool tile_has_sr() { return sizeof(S) == 2; }

template <typename S, int TJ, bool SPLIT, bool SR, bool NT>
static const void *tile_kernel_ptr()
{
    if constexpr (SR && !tile_has_sr<S>()) return nullptr;
    else return reinterpret_cast<const void *>(&a3d_tile_step<S, TJ, SPLIT, SR, NT, false>);
}

// (rows per block, split, non-temporal stores) built
#define SM_TILE_CASES(X) X(1, false, false) X(2, false, false) X(4, false, false) X(8, false, false) \
                         X(1, true, false) X(2, true, false) X(4, true, false) X(8, true, false)     \
                         X(4, false, true) X(8, false, true) X(4, true, true) X(8, true, true)

template <typename S>
int tile_max_blocks_per_cu(int tj, bool split, bool nt, bool sr, bool m16, int *blocks)
{
    if (m16) return tile_max_blocks_per_cu_m16<S>(tj, split, nt, sr, blocks);
    const void *f = nullptr;
#define SM_X(T, P, N)                                                                          \
    if (tj == T && split == P && nt == N)                                                      \
        f = sr ? tile_kernel_ptr<S, T, P, true, N>() : tile_kernel_ptr<S, T, P, false, N>();
    SM_TILE_CASES(SM_X)
#undef SM_X
    if (!f) return SM_ERR_ARG;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64 * tj, 0) != hipSuccess)
        return SM_ERR_HIP;
    *blocks = nb;
    return SM_OK;
}

template <typename S>
int tile_launch(const FusedArgs<S> &a, int tj, bool split, bool nt, bool m16, hipStream_t st)
{
    if (m16) return tile_launch_m16<S>(a, tj, split, nt, st);
    const int blocks = a.nblocks + a.gather_blocks;
    bool done = false;
#define SM_X(T, P, N)                                                                            \
    if (!done && tj == T && split == P && nt == N) {                                             \
        if constexpr (tile_has_sr<S>()) {                                                        \
            if (a.sr) a3d_tile_step<S, T, P, true, N, false><<<blocks, dim3(64, T), 0, st>>>(a); \
            else a3d_tile_step<S, T, P, false, N, false><<<blocks, dim3(64, T), 0, st>>>(a);     \
        } else {                                                                                 \
            a3d_tile_step<S, T, P, false, N, false><<<blocks, dim3(64, T), 0, st>>>(a);          \
        }                                                                                        \
        done = true;                                                                             \
    }
    SM_TILE_CASES(SM_X)
#undef SM_X
    if (!done) return SM_ERR_ARG;
    return hipGetLastError() == hipSuccess ? SM_OK : SM_ERR_HIP;
}

template int tile_launch<float>(const FusedArgs<float> &, int, bool, bool, bool, hipStream_t);
template int tile_launch<__hip_bfloat16>(const FusedArgs<__hip_bfloat16> &, int, bool, bool, bool,
                                         hipStream_t);
template int tile_launch<__half>(const FusedArgs<__half> &, int, bool, bool, bool, hipStream_t);
template int tile_max_blocks_per_cu<float>(int, bool, bool, bool, bool, int *);
template int tile_max_blocks_per_cu<__hip_bfloat16>(int, bool, bool, bool, bool, int *);
template int tile_max_blocks_per_cu<__half>(int, bool, bool, bool, bool, int *);

}  // namespace sm
