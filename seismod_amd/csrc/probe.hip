// probe.hip -- plain streaming kernel of known traffic, used to calibrate
// HBM counters (FETCH_SIZE / WRITE_SIZE) and the achievable bandwidth
// against which the stencil kernels are compared.
#include <algorithm>
#include <vector>

#include "sm_internal.hpp"

namespace sm {

// out[w][c] = sum_r in[r][c] for reads arrays in, writes arrays out
__global__ void __launch_bounds__(256)
probe_stream(const float *__restrict__ in, float *__restrict__ out, size_t n, int reads, int writes)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += stride) {
        float s = 0.f;
        for (int r = 0; r < reads; ++r) s += in[(size_t)r * n + c];
        for (int w = 0; w < writes; ++w) out[(size_t)w * n + c] = s + (float)w;
    }
}

}  // namespace sm

extern "C" int sm_probe_stream(sm_ctx *ctx, size_t n, int reads, int writes, int iters,
                               double *ms_per_iter)
{
    using namespace sm;
    SM_CHECK(ctx && ms_per_iter, SM_ERR_ARG, "sm_probe_stream: NULL argument");
    SM_CHECK(n > 0 && reads >= 1 && reads <= 16 && writes >= 1 && writes <= 16 && iters >= 1,
             SM_ERR_ARG, "sm_probe_stream: bad sizes");
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
    const int grid = std::max(1, dev_cus) * 16;
    probe_stream<<<grid, 256, 0, st>>>((const float *)a.ptr, (float *)b.ptr, n, reads, writes);
    SM_HIP(hipEventRecord(e0, st));
    for (int it = 0; it < iters; ++it)
        probe_stream<<<grid, 256, 0, st>>>((const float *)a.ptr, (float *)b.ptr, n, reads, writes);
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
