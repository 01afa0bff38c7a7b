// Principal submatrix on a set of kept DOFs, renumbered consecutively in
// their order: the free-free block matrix_blocked_ff (src/MatrixUtilityModule.jl:
// 675-690) of a matrix whose numbering lists the kept DOFs first in that
// order, as numberdofs! does for free DOFs (src/FieldModule.jl:343-377).
// A scan of the keep flags gives the new index of every DOF; one thread per
// kept column counts, a scan places, one thread per kept column copies (the
// renumbering is monotone, so rows stay sorted).
#include "common.hpp"

namespace feasm {

namespace {

__global__ __launch_bounds__(256) void k_keep_count(const int64_t *__restrict__ colptr,
                                                    const int32_t *__restrict__ rowval,
                                                    const int32_t *__restrict__ keep,
                                                    const int64_t *__restrict__ newidx, int64_t n,
                                                    int32_t *__restrict__ cnt) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        if (!keep[c]) continue;
        int32_t k = 0;
        for (int64_t p = colptr[c]; p < colptr[c + 1]; p++) k += keep[rowval[p]];
        cnt[newidx[c]] = k;
    }
}

__global__ __launch_bounds__(256) void k_keep_copy(const int64_t *__restrict__ colptr,
                                                   const int32_t *__restrict__ rowval,
                                                   const double *__restrict__ nzval, const int32_t *__restrict__ keep,
                                                   const int64_t *__restrict__ newidx, int64_t n,
                                                   const int64_t *__restrict__ newptr, int32_t *__restrict__ rv2,
                                                   double *__restrict__ nz2) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        if (!keep[c]) continue;
        int64_t q = newptr[newidx[c]];
        for (int64_t p = colptr[c]; p < colptr[c + 1]; p++) {
            const int32_t r = rowval[p];
            if (keep[r]) {
                rv2[q] = (int32_t)newidx[r];
                nz2[q] = nzval[p];
                q++;
            }
        }
    }
}

// entries of column c with row < r (rows are sorted: a binary search)
__device__ __forceinline__ int64_t rows_below(const int64_t *colptr, const int32_t *rowval, int64_t c, int64_t r) {
    int64_t lo = colptr[c], hi = colptr[c + 1];
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)rowval[mid] < r) lo = mid + 1;
        else hi = mid;
    }
    return lo - colptr[c];
}

__global__ __launch_bounds__(256) void k_lead_count(const int64_t *__restrict__ colptr,
                                                    const int32_t *__restrict__ rowval, int64_t nc, int64_t r,
                                                    int32_t *__restrict__ cnt) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x)
        cnt[c] = (int32_t)rows_below(colptr, rowval, c, r);
}

__global__ __launch_bounds__(256) void k_lead_copy(const int64_t *__restrict__ colptr,
                                                   const int32_t *__restrict__ rowval,
                                                   const double *__restrict__ nzval, int64_t nc,
                                                   const int64_t *__restrict__ newptr, int32_t *__restrict__ rv2,
                                                   double *__restrict__ nz2) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = newptr[c + 1] - newptr[c];
        for (int64_t j = 0; j < k; j++) {
            rv2[newptr[c] + j] = rowval[colptr[c] + j];
            nz2[newptr[c] + j] = nzval[colptr[c] + j];
        }
    }
}

}  // namespace

// matrix_blocked_ff(A, r, c) (src/MatrixUtilityModule.jl:675-690): A[1:r, 1:c],
// stored entries (explicit zeros included) kept as they are
feasm_matrix *leading_block(feasm_ctx *ctx, const feasm_matrix *a, int64_t r, int64_t c) {
    FEASM_REQUIRE(r <= a->m, FEASM_EINVAL, "The ff block has too many rows");
    FEASM_REQUIRE(c <= a->n, FEASM_EINVAL, "The ff block has too many columns");
    FEASM_REQUIRE(r >= 0 && c >= 0, FEASM_EINVAL, "negative block size");
    std::unique_ptr<feasm_matrix> M(new feasm_matrix);
    M->ctx = ctx;
    M->m = r;
    M->n = c;
    M->colptr = dev_alloc((c + 1) * sizeof(int64_t));
    if (r == 0 || c == 0 || a->nnz == 0) {  // spzeros(r, c)
        fill_i64(ctx->stream, dp<int64_t>(M->colptr), c + 1, 0);
        FEASM_HIP(hipStreamSynchronize(ctx->stream));
        M->nnz = 0;
        return M.release();
    }
    DevPtr cnt = dev_alloc((c + 1) * sizeof(int32_t));
    FEASM_HIP(hipMemsetAsync(cnt->p, 0, (c + 1) * sizeof(int32_t), ctx->stream));
    hipLaunchKernelGGL(k_lead_count, dim3(grid_for(c, 256)), dim3(256), 0, ctx->stream, dp<int64_t>(a->colptr),
                       dp<int32_t>(a->rowval), c, r, dp<int32_t>(cnt));
    FEASM_HIP(hipGetLastError());
    exclusive_scan_i32_to_i64(ctx, dp<int32_t>(cnt), dp<int64_t>(M->colptr), c + 1);
    int64_t nnz = 0;
    read_back(ctx->stream, &nnz, dp<int64_t>(M->colptr) + c, sizeof(int64_t));
    M->nnz = nnz;
    M->rowval = dev_alloc(std::max<int64_t>(nnz, 1) * sizeof(int32_t));
    M->nzval = dev_alloc(std::max<int64_t>(nnz, 1) * sizeof(double));
    hipLaunchKernelGGL(k_lead_copy, dim3(grid_for(c, 256)), dim3(256), 0, ctx->stream, dp<int64_t>(a->colptr),
                       dp<int32_t>(a->rowval), dp<double>(a->nzval), c, dp<int64_t>(M->colptr),
                       dp<int32_t>(M->rowval), dp<double>(M->nzval));
    FEASM_HIP(hipGetLastError());
    FEASM_HIP(hipStreamSynchronize(ctx->stream));
    return M.release();
}

feasm_matrix *principal_submatrix(feasm_ctx *ctx, const feasm_matrix *a, const uint8_t *keep_host) {
    FEASM_REQUIRE(a->m == a->n, FEASM_EINVAL, "the free block is taken from a square matrix");
    const int64_t n = a->n;
    std::vector<int32_t> kh(n + 1, 0);
    for (int64_t i = 0; i < n; i++) kh[i] = keep_host[i] ? 1 : 0;
    DevPtr keep = dev_alloc((n + 1) * sizeof(int32_t));
    FEASM_HIP(hipMemcpyAsync(keep->p, kh.data(), (n + 1) * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    DevPtr newidx = dev_alloc((n + 1) * sizeof(int64_t));
    exclusive_scan_i32_to_i64(ctx, dp<int32_t>(keep), dp<int64_t>(newidx), n + 1);
    int64_t nf = 0;
    for (int64_t i = 0; i < n; i++) nf += kh[i];
    DevPtr cnt = dev_alloc((nf + 1) * sizeof(int32_t));
    FEASM_HIP(hipMemsetAsync(cnt->p, 0, (nf + 1) * sizeof(int32_t), ctx->stream));
    hipLaunchKernelGGL(k_keep_count, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, dp<int64_t>(a->colptr),
                       dp<int32_t>(a->rowval), dp<int32_t>(keep), dp<int64_t>(newidx), n, dp<int32_t>(cnt));
    FEASM_HIP(hipGetLastError());
    std::unique_ptr<feasm_matrix> M(new feasm_matrix);
    M->ctx = ctx;
    M->m = M->n = nf;
    M->colptr = dev_alloc((nf + 1) * sizeof(int64_t));
    exclusive_scan_i32_to_i64(ctx, dp<int32_t>(cnt), dp<int64_t>(M->colptr), nf + 1);
    int64_t nnz = 0;
    read_back(ctx->stream, &nnz, dp<int64_t>(M->colptr) + nf, sizeof(int64_t));
    M->nnz = nnz;
    M->rowval = dev_alloc(std::max<int64_t>(nnz, 1) * sizeof(int32_t));
    M->nzval = dev_alloc(std::max<int64_t>(nnz, 1) * sizeof(double));
    hipLaunchKernelGGL(k_keep_copy, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, dp<int64_t>(a->colptr),
                       dp<int32_t>(a->rowval), dp<double>(a->nzval), dp<int32_t>(keep), dp<int64_t>(newidx), n,
                       dp<int64_t>(M->colptr), dp<int32_t>(M->rowval), dp<double>(M->nzval));
    FEASM_HIP(hipGetLastError());
    FEASM_HIP(hipStreamSynchronize(ctx->stream));
    return M.release();
}

}  // namespace feasm
