// feasm_matrix_checksum: a 64-bit digest of a device CSC matrix (column
// pointers, row indices and the bits of every value), computed on the device
// so that large results can be compared bitwise without copying them out.
// Each entry is hashed with its position (SplitMix64 finaliser) and the
// hashes are added modulo 2^64, so the digest does not depend on how the
// work is split over threads.
#include "common.hpp"

namespace feasm {

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__global__ void k_zero_u64(unsigned long long *p) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *p = 0ull;
}

__global__ __launch_bounds__(256) void k_digest(const int64_t *__restrict__ colptr, int64_t n,
                                                const int32_t *__restrict__ rowval, const double *__restrict__ nzval,
                                                int64_t nnz, unsigned long long *__restrict__ out) {
    __shared__ unsigned long long part[256];
    uint64_t acc = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz; i += stride) {
        const uint64_t v = (uint64_t)__double_as_longlong(nzval[i]);
        acc += mix64(v ^ mix64((uint64_t)i * 2 + 1)) + mix64(((uint64_t)(uint32_t)rowval[i] << 1) ^ ((uint64_t)i << 33));
    }
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j <= n; j += stride)
        acc += mix64((uint64_t)colptr[j] * 0x2545f4914f6cdd1dull + (uint64_t)j);
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, part[0]);
}

}  // namespace

}  // namespace feasm

using namespace feasm;

extern "C" int feasm_matrix_checksum(feasm_matrix *mat, uint64_t *digest) {
    try {
        FEASM_REQUIRE(mat && digest, FEASM_EINVAL, "null argument");
        feasm_ctx *ctx = mat->ctx;
        FEASM_REQUIRE(ctx, FEASM_ESTATE, "matrix has no device context");
        hipStream_t s = ctx->stream;
        DevPtr acc = dev_alloc(256);
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, dp<unsigned long long>(acc));
        hipLaunchKernelGGL(k_digest, dim3(grid_for(std::max<int64_t>(mat->nnz, mat->n + 1), 256, 4096)), dim3(256), 0,
                           s, dp<int64_t>(mat->colptr), mat->n, dp<int32_t>(mat->rowval), dp<double>(mat->nzval),
                           mat->nnz, dp<unsigned long long>(acc));
        FEASM_HIP(hipGetLastError());
        unsigned long long h = 0;
        read_back(s, &h, acc->p, sizeof(h));
        *digest = (uint64_t)h;
        return FEASM_OK;
    } catch (const feasm::Error &e) {
        if (mat && mat->ctx) mat->ctx->last_error = e.what();
        return e.code;
    }
}
