// Device sparse(I, J, V, m, n): the COO -> CSC stage of makematrix!
// (src/AssemblyModule.jl:304-329), i.e. Julia's SparseArrays.sparse with its
// documented semantics: duplicates combined with + as a left fold in input
// order, explicit zeros kept, rows sorted within every column.
//
//   1. key = col * m + row (64 bit), payload = input position;
//   2. stable LSD radix sort over the key bits actually used -- stability
//      keeps equal keys in input order;
//   3. segment heads where the key changes, scanned to segment numbers;
//   4. one thread per segment folds its values left to right in that order
//      (the same additions, in the same order, as the stdlib), and writes
//      rowval / nzval; column pointers come from the sorted segment columns.
#include "common.hpp"

#include <rocprim/rocprim.hpp>

namespace feasm {

namespace {

__global__ void k_keys(const int32_t *__restrict__ I, const int32_t *__restrict__ J, int64_t nt, int64_t m,
                       uint64_t *__restrict__ keys, uint32_t *__restrict__ idx) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nt; t += (int64_t)gridDim.x * blockDim.x) {
        keys[t] = (uint64_t)J[t] * (uint64_t)m + (uint64_t)I[t];
        idx[t] = (uint32_t)t;
    }
}

__global__ void k_heads(const uint64_t *__restrict__ keys, int64_t nt, int32_t *__restrict__ head) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nt; t += (int64_t)gridDim.x * blockDim.x)
        head[t] = (t == 0 || keys[t] != keys[t - 1]) ? 1 : 0;
}

__global__ void k_seg_start(const int32_t *__restrict__ head, const int64_t *__restrict__ seg, int64_t nt,
                            int64_t *__restrict__ start) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nt; t += (int64_t)gridDim.x * blockDim.x)
        if (head[t]) start[seg[t]] = t;
}

__global__ void k_seg_sum(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ idx,
                          const double *__restrict__ V, const int64_t *__restrict__ start, int64_t nnz, int64_t nt,
                          int64_t m, int32_t *__restrict__ rowval, double *__restrict__ nzval) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < nnz; s += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = start[s], e = (s + 1 < nnz) ? start[s + 1] : nt;
        double v = V[idx[b]];
        for (int64_t t = b + 1; t < e; t++) v = v + V[idx[t]];
        nzval[s] = v;
        rowval[s] = (int32_t)(keys[b] % (uint64_t)m);
    }
}

__global__ void k_colptr(const uint64_t *__restrict__ keys, const int64_t *__restrict__ start, int64_t nnz,
                         int64_t m, int64_t n, int64_t *__restrict__ colptr) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s <= nnz; s += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = (s < nnz) ? (int64_t)(keys[start[s]] / (uint64_t)m) : n;
        const int64_t prev = (s == 0) ? -1 : (int64_t)(keys[start[s - 1]] / (uint64_t)m);
        for (int64_t j = prev + 1; j <= c; j++) colptr[j] = s;
    }
}

int bits_for(uint64_t v) {
    int b = 0;
    while (b < 64 && (v >> b) != 0) b++;
    return b < 1 ? 1 : b;
}

}  // namespace

feasm_matrix *sparse_device(feasm_ctx *ctx, int64_t nt, const int32_t *I, const int32_t *J, const double *V,
                            int64_t m, int64_t n) {
    FEASM_REQUIRE(nt >= 0 && m >= 0 && n >= 0, FEASM_EINVAL, "negative size");
    FEASM_REQUIRE(nt < (1ll << 32), FEASM_EUNSUPPORTED, "more than 2^32 triples");
    std::unique_ptr<feasm_matrix> M(new feasm_matrix);
    M->ctx = ctx;
    M->m = m;
    M->n = n;
    M->colptr = dev_alloc((n + 1) * sizeof(int64_t));
    if (nt == 0) {
        FEASM_HIP(hipMemsetAsync(M->colptr->p, 0, (n + 1) * sizeof(int64_t), ctx->stream));
        FEASM_HIP(hipStreamSynchronize(ctx->stream));
        M->nnz = 0;
        return M.release();
    }
    DevPtr keys = dev_alloc(nt * sizeof(uint64_t)), keys2 = dev_alloc(nt * sizeof(uint64_t));
    DevPtr idx = dev_alloc(nt * sizeof(uint32_t)), idx2 = dev_alloc(nt * sizeof(uint32_t));
    hipLaunchKernelGGL(k_keys, dim3(grid_for(nt, 256)), dim3(256), 0, ctx->stream, I, J, nt, m, dp<uint64_t>(keys),
                       dp<uint32_t>(idx));
    FEASM_HIP(hipGetLastError());
    const unsigned end_bit = (unsigned)bits_for((uint64_t)m * (uint64_t)(n > 0 ? n : 1));
    size_t bytes = 0;
    FEASM_HIP(rocprim::radix_sort_pairs(nullptr, bytes, dp<uint64_t>(keys), dp<uint64_t>(keys2), dp<uint32_t>(idx),
                                        dp<uint32_t>(idx2), (size_t)nt, 0u, end_bit, ctx->stream));
    void *tmp = ctx->scratch_get(bytes);
    FEASM_HIP(rocprim::radix_sort_pairs(tmp, bytes, dp<uint64_t>(keys), dp<uint64_t>(keys2), dp<uint32_t>(idx),
                                        dp<uint32_t>(idx2), (size_t)nt, 0u, end_bit, ctx->stream));
    DevPtr head = dev_alloc(nt * sizeof(int32_t));
    hipLaunchKernelGGL(k_heads, dim3(grid_for(nt, 256)), dim3(256), 0, ctx->stream, dp<uint64_t>(keys2), nt,
                       dp<int32_t>(head));
    DevPtr seg = dev_alloc(nt * sizeof(int64_t));
    exclusive_scan_i32_to_i64(ctx, dp<int32_t>(head), dp<int64_t>(seg), nt);
    int64_t last_seg = 0;
    int32_t last_head = 0;
    FEASM_HIP(hipMemcpyAsync(&last_seg, dp<int64_t>(seg) + nt - 1, sizeof(int64_t), hipMemcpyDeviceToHost,
                             ctx->stream));
    FEASM_HIP(hipMemcpyAsync(&last_head, dp<int32_t>(head) + nt - 1, sizeof(int32_t), hipMemcpyDeviceToHost,
                             ctx->stream));
    FEASM_HIP(hipStreamSynchronize(ctx->stream));
    const int64_t nnz = last_seg + last_head;
    DevPtr start = dev_alloc(nnz * sizeof(int64_t));
    hipLaunchKernelGGL(k_seg_start, dim3(grid_for(nt, 256)), dim3(256), 0, ctx->stream, dp<int32_t>(head),
                       dp<int64_t>(seg), nt, dp<int64_t>(start));
    M->nnz = nnz;
    M->rowval = dev_alloc(nnz * sizeof(int32_t));
    M->nzval = dev_alloc(nnz * sizeof(double));
    hipLaunchKernelGGL(k_seg_sum, dim3(grid_for(nnz, 256)), dim3(256), 0, ctx->stream, dp<uint64_t>(keys2),
                       dp<uint32_t>(idx2), V, dp<int64_t>(start), nnz, nt, m, dp<int32_t>(M->rowval),
                       dp<double>(M->nzval));
    hipLaunchKernelGGL(k_colptr, dim3(grid_for(nnz + 1, 256)), dim3(256), 0, ctx->stream, dp<uint64_t>(keys2),
                       dp<int64_t>(start), nnz, m, n, dp<int64_t>(M->colptr));
    FEASM_HIP(hipGetLastError());
    FEASM_HIP(hipStreamSynchronize(ctx->stream));
    return M.release();
}

}  // namespace feasm
