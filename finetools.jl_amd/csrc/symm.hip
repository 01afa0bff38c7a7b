// makematrix! of SysmatAssemblerSparseSymm (src/AssemblyModule.jl:551-583):
// given S = sparse() of the buffered lower-triangle triples, the result is
// S + transpose(S) with its diagonal halved.  Each stored (i, j, v) of S is
// listed twice -- as (i, j, v) and as (j, i, v) -- and the device sparse()
// sums the list: an off-diagonal position receives at most one value from S
// and one from its transpose (a + b, which does not depend on the order), a
// diagonal position S(j,j) twice (exactly 2 S(j,j)).  SparseArrays' `+`
// stores no zero result, so exact zeros are dropped (before the halving,
// which leaves nonzero values nonzero); then the diagonal is halved.
#include "common.hpp"

namespace feasm {

namespace {

__global__ __launch_bounds__(256) void k_expand_both(const int64_t *__restrict__ colptr,
                                                     const int32_t *__restrict__ rowval,
                                                     const double *__restrict__ nzval, int64_t n, int64_t nnz,
                                                     int32_t *__restrict__ I, int32_t *__restrict__ J,
                                                     double *__restrict__ V) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x)
        for (int64_t p = colptr[c]; p < colptr[c + 1]; p++) {
            const int32_t r = rowval[p];
            const double v = nzval[p];
            I[p] = r;
            J[p] = (int32_t)c;
            V[p] = v;
            I[nnz + p] = (int32_t)c;
            J[nnz + p] = r;
            V[nnz + p] = v;
        }
}

__global__ __launch_bounds__(256) void k_halve_diagonal(const int64_t *__restrict__ colptr,
                                                        const int32_t *__restrict__ rowval, int64_t n,
                                                        double *__restrict__ nzval) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x)
        for (int64_t p = colptr[c]; p < colptr[c + 1]; p++)
            if (rowval[p] == (int32_t)c) nzval[p] *= 0.5;
}

}  // namespace

feasm_matrix *symm_plus_transpose(feasm_ctx *ctx, const feasm_matrix *S) {
    FEASM_REQUIRE(S->m == S->n, FEASM_EINVAL, "Row and column info do not agree");
    const int64_t n = S->n, nnz = S->nnz;
    DevPtr I = dev_alloc(std::max<size_t>(2 * nnz * sizeof(int32_t), 4));
    DevPtr J = dev_alloc(std::max<size_t>(2 * nnz * sizeof(int32_t), 4));
    DevPtr V = dev_alloc(std::max<size_t>(2 * nnz * sizeof(double), 8));
    if (nnz) {
        hipLaunchKernelGGL(k_expand_both, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, dp<int64_t>(S->colptr),
                           dp<int32_t>(S->rowval), dp<double>(S->nzval), n, nnz, dp<int32_t>(I), dp<int32_t>(J),
                           dp<double>(V));
        FEASM_HIP(hipGetLastError());
    }
    std::unique_ptr<feasm_matrix> T(sparse_device(ctx, 2 * nnz, dp<int32_t>(I), dp<int32_t>(J), dp<double>(V), n, n));
    std::unique_ptr<feasm_matrix> R(drop_zeros(ctx, T.get()));
    if (R->nnz) {
        hipLaunchKernelGGL(k_halve_diagonal, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream,
                           dp<int64_t>(R->colptr), dp<int32_t>(R->rowval), n, dp<double>(R->nzval));
        FEASM_HIP(hipGetLastError());
    }
    FEASM_HIP(hipStreamSynchronize(ctx->stream));
    return R.release();
}

}  // namespace feasm
