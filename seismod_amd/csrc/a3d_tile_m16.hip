// a3d_tile_m16.hip -- launches of the LDS-staged 3D step (a3d_tile.hpp)
// whose material (rho, K*dt) is stored as FP16 with one power-of-two scale
// per field: 4 instead of 8 bytes of material per cell.  Built for the
// 16-bit state types only, 4 or 8 rows per block, plain and split CPML
// forms, cached and non-temporal state stores, round-to-nearest stores.
#include "a3d_tile.hpp"

namespace sm {

#define SM_TILE_M16_CASES(X) X(4, false, false) X(8, false, false) X(4, true, false) X(8, true, false) \
                             X(4, false, true) X(8, false, true) X(4, true, true) X(8, true, true)

template <typename S>
int tile_max_blocks_per_cu_m16(int tj, bool split, bool nt, bool sr, int *blocks)
{
    if constexpr (sizeof(S) != 2) {
        return SM_ERR_ARG;
    } else {
        if (sr) return SM_ERR_ARG;
        const void *f = nullptr;
#define SM_X(T, P, N)                                                                    \
        if (tj == T && split == P && nt == N)                                            \
            f = reinterpret_cast<const void *>(&a3d_tile_step<S, T, P, false, N, true>);
        SM_TILE_M16_CASES(SM_X)
#undef SM_X
        if (!f) return SM_ERR_ARG;
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64 * tj, 0) != hipSuccess)
            return SM_ERR_HIP;
        *blocks = nb;
        return SM_OK;
    }
}

template <typename S>
int tile_launch_m16(const FusedArgs<S> &a, int tj, bool split, bool nt, hipStream_t st)
{
    if constexpr (sizeof(S) != 2) {
        return SM_ERR_ARG;
    } else {
        if (a.sr) return SM_ERR_ARG;
        const int blocks = a.nblocks + a.gather_blocks;
        bool done = false;
#define SM_X(T, P, N)                                                                    \
        if (!done && tj == T && split == P && nt == N) {                                 \
            a3d_tile_step<S, T, P, false, N, true><<<blocks, dim3(64, T), 0, st>>>(a);   \
            done = true;                                                                 \
        }
        SM_TILE_M16_CASES(SM_X)
#undef SM_X
        if (!done) return SM_ERR_ARG;
        return hipGetLastError() == hipSuccess ? SM_OK : SM_ERR_HIP;
    }
}

template int tile_launch_m16<float>(const FusedArgs<float> &, int, bool, bool, hipStream_t);
template int tile_launch_m16<__hip_bfloat16>(const FusedArgs<__hip_bfloat16> &, int, bool, bool,
                                             hipStream_t);
template int tile_launch_m16<__half>(const FusedArgs<__half> &, int, bool, bool, hipStream_t);
template int tile_max_blocks_per_cu_m16<float>(int, bool, bool, bool, int *);
template int tile_max_blocks_per_cu_m16<__hip_bfloat16>(int, bool, bool, bool, int *);
template int tile_max_blocks_per_cu_m16<__half>(int, bool, bool, bool, int *);

}  // namespace sm
