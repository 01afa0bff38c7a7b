// Measurement helper: a plain streaming copy of known size, used to scale the
// FETCH_SIZE / WRITE_SIZE hardware counters of the assembly kernels to HBM
// bytes (a copy of N bytes must read N and write N).
#include "common.hpp"

namespace feasm {
namespace {
__global__ __launch_bounds__(256) void k_copy_calib(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                    int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
}  // namespace
}  // namespace feasm

extern "C" int feasm_copy_calibration(feasm_ctx *ctx, int64_t nbytes, int repeats, double *ms_per_copy) {
    try {
        FEASM_REQUIRE(ctx && nbytes >= 16 && repeats >= 1, FEASM_EINVAL, "bad argument");
        const int64_t n = nbytes / 16;
        feasm::DevPtr a = feasm::dev_alloc(n * 16), b = feasm::dev_alloc(n * 16);
        FEASM_HIP(hipMemsetAsync(a->p, 0, n * 16, ctx->stream));
        hipEvent_t e0, e1;
        FEASM_HIP(hipEventCreate(&e0));
        FEASM_HIP(hipEventCreate(&e1));
        const unsigned grid = feasm::grid_for(n, 256, 1u << 16);
        FEASM_HIP(hipEventRecord(e0, ctx->stream));
        for (int r = 0; r < repeats; r++)
            hipLaunchKernelGGL(feasm::k_copy_calib, dim3(grid), dim3(256), 0, ctx->stream,
                               static_cast<const double2 *>(a->p), static_cast<double2 *>(b->p), n);
        FEASM_HIP(hipEventRecord(e1, ctx->stream));
        FEASM_HIP(hipEventSynchronize(e1));
        float ms = 0;
        FEASM_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (ms_per_copy) *ms_per_copy = ms / repeats;
        return FEASM_OK;
    } catch (const feasm::Error &e) {
        ctx->last_error = e.what();
        return e.code;
    }
}
