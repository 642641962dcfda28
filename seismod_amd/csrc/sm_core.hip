// sm_core.hip -- context, errors and the FP64 host set-up shared by both
// solvers (Ricker wavelet, step count, record predicate).
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>

#include "e2d_fused.hpp"
#include "sm_internal.hpp"

namespace sm {

static thread_local char g_last_error[512] = "";

int fail(int status, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return status;
}

int set_device(const sm_ctx *ctx)
{
    SM_HIP(hipSetDevice(ctx->device));
    return SM_OK;
}

int dalloc(sm_dbuf &b, size_t bytes)
{
    b.bytes = bytes;
    b.ptr = nullptr;
    if (bytes == 0) return SM_OK;
    hipError_t e = hipMalloc(&b.ptr, bytes);
    if (e != hipSuccess) {
        b.ptr = nullptr;
        b.bytes = 0;
        return fail(SM_ERR_NOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    }
    return SM_OK;
}

void dfree(sm_dbuf &b)
{
    if (b.ptr) (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
}

}  // namespace sm

extern "C" {

int sm_abi_version(void) { return SM_ABI_VERSION; }

const char *sm_strerror(int status)
{
    switch (status) {
    case SM_OK: return "ok";
    case SM_ERR_ARG: return "invalid argument";
    case SM_ERR_HIP: return "HIP runtime error";
    case SM_ERR_CFL: return "time step too large (stability limit)";
    case SM_ERR_NOMEM: return "out of memory";
    case SM_ERR_STATE: return "invalid solver state";
    case SM_ERR_NODEV: return "no usable GPU";
    default: return "unknown status";
    }
}

const char *sm_last_error(void) { return sm::g_last_error; }

int sm_ctx_create(int device, sm_ctx **out)
{
    SM_CHECK(out != nullptr, SM_ERR_ARG, "sm_ctx_create: out is NULL");
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        return sm::fail(SM_ERR_NODEV, "no HIP device (%s)", hipGetErrorString(e));
    SM_CHECK(device >= 0 && device < count, SM_ERR_ARG, "device %d out of range (%d devices)",
             device, count);
    SM_HIP(hipSetDevice(device));
    // per-device address of the allocation-free batches' scratch array
    const int rc = sm::e2d_scratch_init(device);
    if (rc != SM_OK) return rc;
    hipStream_t st;
    SM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    sm_ctx *c = new sm_ctx;
    c->device = device;
    c->stream = st;
    *out = c;
    return SM_OK;
}

int sm_ctx_destroy(sm_ctx *ctx)
{
    if (!ctx) return SM_OK;
    SM_HIP(hipSetDevice(ctx->device));
    SM_HIP(hipStreamSynchronize(ctx->stream));
    SM_HIP(hipStreamDestroy(ctx->stream));
    delete ctx;
    return SM_OK;
}

int sm_ctx_sync(sm_ctx *ctx)
{
    SM_CHECK(ctx, SM_ERR_ARG, "sm_ctx_sync: NULL context");
    SM_HIP(hipSetDevice(ctx->device));
    SM_HIP(hipStreamSynchronize(ctx->stream));
    return SM_OK;
}

int sm_ctx_mem_info(sm_ctx *ctx, size_t *free_bytes, size_t *total_bytes)
{
    SM_CHECK(ctx && free_bytes && total_bytes, SM_ERR_ARG, "sm_ctx_mem_info: NULL argument");
    SM_HIP(hipSetDevice(ctx->device));
    SM_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return SM_OK;
}

size_t sm_ricker_length(double f, double dt)
{
    // src.c:50-51
    const double tcut = 1.0 / f;
    return (size_t)std::ceil(tcut * 2.0 / dt);
}

int sm_compute_src_fct(size_t n, const double *A, const double *f, double dt,
                       const size_t *length, double *fct)
{
    SM_CHECK(n > 0 && A && f && length && fct, SM_ERR_ARG, "sm_compute_src_fct: bad arguments");
    SM_CHECK(dt > 0.0, SM_ERR_ARG, "sm_compute_src_fct: dt must be > 0");
    // src.c:45-64.  Each source's pass takes its maximum over the samples of
    // source 0 and rescales source 0 by A_s/max; later sources stay raw.
    const double pi = 4.0 * std::atan(1.0);
    const double pi2 = pi * pi;
    double *w0 = fct;
    const size_t len0 = length[0];
    size_t off = 0;
    for (size_t s = 0; s < n; ++s) {
        SM_CHECK(f[s] > 0.0, SM_ERR_ARG, "source %zu: frequency must be > 0", s);
        double *w = fct + off;
        const double fs = f[s], As = A[s];
        const double tcut = 1.0 / fs;
        double peak = 0.0;
        for (size_t i = 0; i < length[s]; ++i) {
            const double tau = i * dt - tcut;
            const double K = pi2 * fs * fs * tau * tau;
            w[i] = As * 2. * pi2 * fs * fs * (1.0 - 2.0 * K) * std::exp(-K);
            if (i < len0) {  // the reference reads past source 0 here (undefined)
                const double m = std::fabs(w0[i]);
                peak = peak > m ? peak : m;
            }
        }
        for (size_t i = 0; i < len0; ++i) w0[i] *= As / peak;
        off += length[s];
    }
    return SM_OK;
}

size_t sm_nsteps(double duration, double dt)
{
    size_t nsteps = duration / dt;          // e_iso.c:575
    if (nsteps * dt < duration) nsteps++;   // e_iso.c:576
    return nsteps;
}

int sm_record_due(size_t it, double dt, double t0, double rdt)
{
    // e_iso.c:591, 643-644, 657-658
    const double t = dt * (it + 1);
    const double dt2 = t - t0 + 1.e-15;
    const double rem = std::fmod(1.e6 * dt2, 1.e6 * rdt);
    return (std::fabs(dt2) < 0.99 * dt) || (t > t0 && rdt > 0.0 && std::fabs(rem) < 1.e-8);
}

}  // extern "C"
