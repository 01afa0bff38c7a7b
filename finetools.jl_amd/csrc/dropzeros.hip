// dropzeros: a CSC without its stored exact zeros (SparseArrays.dropzeros
// semantics).  The symmetric assembler's makematrix! forms S + transpose(S)
// (src/AssemblyModule.jl:567-583), and sparse + stores no zero result, so its
// pattern is the full symmetric matrix's minus the entries that sum to zero.
// One thread per column counts, a scan places, one thread per column copies
// (rows stay sorted).
#include "common.hpp"

namespace feasm {

namespace {

__global__ __launch_bounds__(256) void k_count_nonzero(const int64_t *__restrict__ colptr,
                                                       const double *__restrict__ nzval, int64_t n,
                                                       int32_t *__restrict__ cnt) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        int32_t k = 0;
        for (int64_t p = colptr[c]; p < colptr[c + 1]; p++) k += (nzval[p] != 0.0) ? 1 : 0;
        cnt[c] = k;
    }
}

__global__ __launch_bounds__(256) void k_copy_nonzero(const int64_t *__restrict__ colptr,
                                                      const int32_t *__restrict__ rowval,
                                                      const double *__restrict__ nzval, int64_t n,
                                                      const int64_t *__restrict__ newptr, int32_t *__restrict__ rv2,
                                                      double *__restrict__ nz2) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        int64_t q = newptr[c];
        for (int64_t p = colptr[c]; p < colptr[c + 1]; p++)
            if (nzval[p] != 0.0) {
                rv2[q] = rowval[p];
                nz2[q] = nzval[p];
                q++;
            }
    }
}

}  // namespace

feasm_matrix *drop_zeros(feasm_ctx *ctx, const feasm_matrix *a) {
    const int64_t n = a->n;
    DevPtr cnt = dev_alloc((n + 1) * sizeof(int32_t));
    FEASM_HIP(hipMemsetAsync(cnt->p, 0, (n + 1) * sizeof(int32_t), ctx->stream));
    hipLaunchKernelGGL(k_count_nonzero, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, dp<int64_t>(a->colptr),
                       dp<double>(a->nzval), n, dp<int32_t>(cnt));
    FEASM_HIP(hipGetLastError());
    std::unique_ptr<feasm_matrix> M(new feasm_matrix);
    M->ctx = ctx;
    M->m = a->m;
    M->n = n;
    M->colptr = dev_alloc((n + 1) * sizeof(int64_t));
    exclusive_scan_i32_to_i64(ctx, dp<int32_t>(cnt), dp<int64_t>(M->colptr), n + 1);
    int64_t nnz = 0;
    read_back(ctx->stream, &nnz, dp<int64_t>(M->colptr) + n, sizeof(int64_t));
    M->nnz = nnz;
    M->rowval = dev_alloc(std::max<int64_t>(nnz, 1) * sizeof(int32_t));
    M->nzval = dev_alloc(std::max<int64_t>(nnz, 1) * sizeof(double));
    hipLaunchKernelGGL(k_copy_nonzero, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, dp<int64_t>(a->colptr),
                       dp<int32_t>(a->rowval), dp<double>(a->nzval), n, dp<int64_t>(M->colptr),
                       dp<int32_t>(M->rowval), dp<double>(M->nzval));
    FEASM_HIP(hipGetLastError());
    FEASM_HIP(hipStreamSynchronize(ctx->stream));
    return M.release();
}

}  // namespace feasm
