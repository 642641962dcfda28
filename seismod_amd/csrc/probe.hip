// probe.hip -- plain streaming kernel of known traffic, used to calibrate
// HBM counters (FETCH_SIZE / WRITE_SIZE) and the achievable bandwidth
// against which the stencil kernels are compared.
#include <algorithm>
#include <vector>

#include "sm_internal.hpp"

namespace sm {

// out[w][c] = sum_r in[r][c] for R arrays in, W arrays out; the array
// counts are compile-time so every load of an element is issued before the
// first is used (no latency chain between streams).  Each lane moves V
// consecutive floats per array (V = 1, 2 or 4: dword, dwordx2, dwordx4);
// NT marks the accesses non-temporal (streamed past the caches).
template <int V> struct VecT;
template <> struct VecT<1> { using T = float; };
template <> struct VecT<2> { using T = float __attribute__((ext_vector_type(2))); };
template <> struct VecT<4> { using T = float __attribute__((ext_vector_type(4))); };

template <int R, int W, int V, bool NT>
__global__ void __launch_bounds__(256)
probe_stream(const float *__restrict__ in, float *__restrict__ out, size_t n)
{
    using T = typename VecT<V>::T;
    const size_t nv = n / V;
    const T *vin = (const T *)in;
    T *vout = (T *)out;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < nv; c += stride) {
        T v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if constexpr (NT) v[r] = __builtin_nontemporal_load(vin + (size_t)r * nv + c);
            else v[r] = vin[(size_t)r * nv + c];
        }
        T s = v[0];
#pragma unroll
        for (int r = 1; r < R; ++r) s = s + v[r];
#pragma unroll
        for (int w = 0; w < W; ++w) {
            if constexpr (NT) __builtin_nontemporal_store(s, vout + (size_t)w * nv + c);
            else vout[(size_t)w * nv + c] = s;
        }
    }
}

using ProbeFn = void (*)(const float *, float *, size_t);

template <int R, int W>
ProbeFn probe_pick(int vec, bool nt)
{
    if (nt) return vec == 4 ? probe_stream<R, W, 4, true> : vec == 2 ? probe_stream<R, W, 2, true>
                                                                     : probe_stream<R, W, 1, true>;
    return vec == 4 ? probe_stream<R, W, 4, false> : vec == 2 ? probe_stream<R, W, 2, false>
                                                              : probe_stream<R, W, 1, false>;
}

ProbeFn probe_kernel(int reads, int writes, int vec, bool nt)
{
#define SM_PROBE(r, w)                                                              \
    if (reads == r && writes == w) return probe_pick<r, w>(vec, nt);
    SM_PROBE(1, 1) SM_PROBE(2, 1) SM_PROBE(3, 1) SM_PROBE(4, 2) SM_PROBE(6, 4) SM_PROBE(8, 4)
    SM_PROBE(10, 4) SM_PROBE(12, 8) SM_PROBE(1, 2) SM_PROBE(5, 5)
#undef SM_PROBE
    return nullptr;
}

}  // namespace sm

extern "C" int sm_probe_stream(sm_ctx *ctx, size_t n, int reads, int writes, int vec,
                               int iters, double *ms_per_iter)
{
    using namespace sm;
    SM_CHECK(ctx && ms_per_iter, SM_ERR_ARG, "sm_probe_stream: NULL argument");
    SM_CHECK(n > 0 && iters >= 1, SM_ERR_ARG, "sm_probe_stream: bad sizes");
    const bool nt = (vec & SM_PROBE_NT) != 0, tiles = (vec & SM_PROBE_TILES) != 0;
    vec &= ~(SM_PROBE_NT | SM_PROBE_TILES);
    SM_CHECK(vec == 1 || vec == 2 || vec == 4, SM_ERR_ARG, "sm_probe_stream: vec must be 1, 2 or 4");
    SM_CHECK(n % vec == 0, SM_ERR_ARG, "sm_probe_stream: n must be a multiple of vec");
    const ProbeFn fn = probe_kernel(reads, writes, vec, nt);
    SM_CHECK(fn, SM_ERR_ARG,
             "sm_probe_stream: (reads, writes) = (%d, %d) not built; available (1,1) (2,1) (3,1) "
             "(4,2) (6,4) (8,4) (10,4) (12,8) (1,2) (5,5)", reads, writes);
    SM_CHECK(set_device(ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    sm_dbuf a, b;
    int rc = dalloc(a, (size_t)reads * n * sizeof(float));
    if (rc == SM_OK) rc = dalloc(b, (size_t)writes * n * sizeof(float));
    if (rc != SM_OK) { dfree(a); dfree(b); return rc; }
    hipStream_t st = ctx->stream;
    hipEvent_t e0, e1;
    SM_HIP(hipMemsetAsync(a.ptr, 0, a.bytes, st));
    SM_HIP(hipEventCreate(&e0));
    SM_HIP(hipEventCreate(&e1));
    int dev_cus = 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess) dev_cus = prop.multiProcessorCount;
    // persistent grid (16 workgroups per CU, each striding over the arrays)
    // or one 256-element tile per workgroup
    const size_t nv = n / vec;
    const int grid = tiles ? (int)std::min<size_t>((nv + 255) / 256, 0x7FFFFFFF)
                           : std::max(1, dev_cus) * 16;
    fn<<<grid, 256, 0, st>>>((const float *)a.ptr, (float *)b.ptr, n);
    SM_HIP(hipEventRecord(e0, st));
    for (int it = 0; it < iters; ++it) fn<<<grid, 256, 0, st>>>((const float *)a.ptr, (float *)b.ptr, n);
    SM_HIP(hipEventRecord(e1, st));
    SM_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    SM_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dfree(a);
    dfree(b);
    *ms_per_iter = ms / iters;
    return SM_OK;
}
