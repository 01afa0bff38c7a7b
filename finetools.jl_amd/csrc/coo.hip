// Device sparse(I, J, V, m, n): the COO -> CSC stage of makematrix!
// (src/AssemblyModule.jl:304-329), i.e. Julia's SparseArrays.sparse with its
// documented semantics: duplicates combined with + as a left fold in input
// order, explicit zeros kept, rows sorted within every column.
//
//   1. key = col * m + row (64 bit), payload = input position; the same pass
//      flags any index outside [0, m) x [0, n) (the stdlib raises an error);
//   2. stable LSD radix sort over the key bits actually used -- stability
//      keeps equal keys in input order;
//   3. segment heads where the key changes, scanned to segment numbers;
//   4. one thread per segment folds its values left to right in that order
//      (the same additions, in the same order, as the stdlib), and writes
//      rowval / nzval; column pointers come from the sorted segment columns.
//
// Steps 1-3 (phase A) end with one read-back of the entry count and the range
// flag; step 4 (phase B) then writes the outputs.
#include "common.hpp"

#include <rocprim/rocprim.hpp>

namespace feasm {

namespace {

__global__ void k_zero_flag(unsigned int *__restrict__ flag) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *flag = 0u;
}

__global__ void k_keys(const int32_t *__restrict__ I, const int32_t *__restrict__ J, int64_t nt, int64_t m, int64_t n,
                       uint64_t *__restrict__ keys, uint32_t *__restrict__ idx, unsigned int *__restrict__ bad) {
    bool out = false;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nt; t += (int64_t)gridDim.x * blockDim.x) {
        const int32_t i = I[t], j = J[t];
        out |= (i < 0) | ((int64_t)i >= m) | (j < 0) | ((int64_t)j >= n);
        keys[t] = (uint64_t)(uint32_t)j * (uint64_t)m + (uint64_t)(uint32_t)i;
        idx[t] = (uint32_t)t;
    }
    if (out) atomicOr(bad, 1u);
}

__global__ void k_heads(const uint64_t *__restrict__ keys, int64_t nt, int32_t *__restrict__ head) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nt; t += (int64_t)gridDim.x * blockDim.x)
        head[t] = (t == 0 || keys[t] != keys[t - 1]) ? 1 : 0;
}

// nnz = last segment number + last head; copied back with the range flag
__global__ void k_count(const int32_t *__restrict__ head, const int64_t *__restrict__ seg, int64_t nt,
                        const unsigned int *__restrict__ bad, int64_t *__restrict__ out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        out[0] = seg[nt - 1] + head[nt - 1];
        out[1] = (int64_t)*bad;
    }
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

// colptr of a matrix with no entries (a fill kernel rather than a memset)
__global__ void k_fill_i64(int64_t *__restrict__ p, int64_t n, int64_t v) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        p[t] = v;
}

}  // namespace

int bits_for(uint64_t v) {
    int b = 0;
    while (b < 64 && (v >> b) != 0) b++;
    return b < 1 ? 1 : b;
}

size_t sort_tmp_bytes(int64_t nt, unsigned end_bit, hipStream_t s) {
    size_t a = 0, b = 0;
    rocprim::double_buffer<uint64_t> kb(nullptr, nullptr);
    rocprim::double_buffer<uint32_t> vb(nullptr, nullptr);
    FEASM_HIP(rocprim::radix_sort_pairs(nullptr, a, kb, vb, (size_t)nt, 0u, end_bit, s));
    FEASM_HIP(rocprim::exclusive_scan(nullptr, b, (const int32_t *)nullptr, (int64_t *)nullptr, (int64_t)0,
                                      (size_t)nt, rocprim::plus<int64_t>(), s));
    return std::max(a, b);
}

// Phase A: keys + range check, stable sort, segment heads and numbers; one
// read-back.  Returns nnz; FEASM_EDOF when an index is out of range.
int64_t sparse_phase_a(hipStream_t s, int64_t nt, const int32_t *I, const int32_t *J, int64_t m, int64_t n,
                       unsigned end_bit, SparseWork &w) {
    hipLaunchKernelGGL(k_zero_flag, dim3(1), dim3(64), 0, s, w.flag);
    hipLaunchKernelGGL(k_keys, dim3(grid_for(nt, 256)), dim3(256), 0, s, I, J, nt, m, n, w.keys, w.idx, w.flag);
    FEASM_HIP(hipGetLastError());
    rocprim::double_buffer<uint64_t> kb(w.keys, w.keys_alt);
    rocprim::double_buffer<uint32_t> vb(w.idx, w.idx_alt);
    size_t bytes = w.tmp_bytes;
    FEASM_HIP(rocprim::radix_sort_pairs(w.tmp, bytes, kb, vb, (size_t)nt, 0u, end_bit, s));
    w.skeys = kb.current();
    w.sidx = vb.current();
    w.head = reinterpret_cast<int32_t *>(vb.alternate());
    w.seg = reinterpret_cast<int64_t *>(kb.alternate());
    hipLaunchKernelGGL(k_heads, dim3(grid_for(nt, 256)), dim3(256), 0, s, w.skeys, nt, w.head);
    bytes = w.tmp_bytes;
    FEASM_HIP(rocprim::exclusive_scan(w.tmp, bytes, (const int32_t *)w.head, w.seg, (int64_t)0, (size_t)nt,
                                      rocprim::plus<int64_t>(), s));
    hipLaunchKernelGGL(k_count, dim3(1), dim3(64), 0, s, (const int32_t *)w.head, (const int64_t *)w.seg, nt,
                       (const unsigned int *)w.flag, w.count);
    FEASM_HIP(hipGetLastError());
    int64_t hc[2] = {0, 0};
    FEASM_HIP(hipMemcpyAsync(hc, w.count, sizeof(hc), hipMemcpyDeviceToHost, s));
    FEASM_HIP(hipStreamSynchronize(s));
    FEASM_REQUIRE(hc[1] == 0, FEASM_EDOF, "sparse: row or column index out of range");
    return hc[0];
}

// Phase B: segment starts, ordered sums, column pointers.
void sparse_phase_b(hipStream_t s, int64_t nt, const double *V, int64_t m, int64_t n, int64_t nnz,
                    const SparseWork &w, int64_t *colptr, int32_t *rowval, double *nzval) {
    hipLaunchKernelGGL(k_seg_start, dim3(grid_for(nt, 256)), dim3(256), 0, s, (const int32_t *)w.head,
                       (const int64_t *)w.seg, nt, w.start);
    hipLaunchKernelGGL(k_seg_sum, dim3(grid_for(nnz, 256)), dim3(256), 0, s, w.skeys, w.sidx, V,
                       (const int64_t *)w.start, nnz, nt, m, rowval, nzval);
    hipLaunchKernelGGL(k_colptr, dim3(grid_for(nnz + 1, 256)), dim3(256), 0, s, w.skeys, (const int64_t *)w.start,
                       nnz, m, n, colptr);
    FEASM_HIP(hipGetLastError());
}

void fill_i64(hipStream_t s, int64_t *p, int64_t n, int64_t v) {
    hipLaunchKernelGGL(k_fill_i64, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, v);
    FEASM_HIP(hipGetLastError());
}

void check_sparse_args(int64_t nt, int64_t m, int64_t n) {
    FEASM_REQUIRE(nt >= 0 && m >= 0 && n >= 0, FEASM_EINVAL, "negative size");
    FEASM_REQUIRE(nt < (1ll << 32), FEASM_EUNSUPPORTED, "more than 2^32 triples");
    FEASM_REQUIRE(m < (1ll << 31) && n < (1ll << 31), FEASM_EUNSUPPORTED, "more than 2^31 rows or columns");
    FEASM_REQUIRE(nt == 0 || (m >= 1 && n >= 1), FEASM_EDOF, "sparse: row or column index out of range");
}

feasm_matrix *sparse_device(feasm_ctx *ctx, int64_t nt, const int32_t *I, const int32_t *J, const double *V,
                            int64_t m, int64_t n) {
    check_sparse_args(nt, m, n);
    hipStream_t s = ctx->stream;
    std::unique_ptr<feasm_matrix> M(new feasm_matrix);
    M->ctx = ctx;
    M->m = m;
    M->n = n;
    M->colptr = dev_alloc((n + 1) * sizeof(int64_t));
    if (nt == 0) {
        fill_i64(s, dp<int64_t>(M->colptr), n + 1, 0);
        FEASM_HIP(hipStreamSynchronize(s));
        M->nnz = 0;
        return M.release();
    }
    const unsigned end_bit = (unsigned)bits_for((uint64_t)m * (uint64_t)n);
    SparseWork w;
    DevPtr keys = dev_alloc(nt * sizeof(uint64_t)), keys2 = dev_alloc(nt * sizeof(uint64_t));
    DevPtr idx = dev_alloc(nt * sizeof(uint32_t)), idx2 = dev_alloc(nt * sizeof(uint32_t));
    DevPtr small = dev_alloc(256);
    w.keys = dp<uint64_t>(keys);
    w.keys_alt = dp<uint64_t>(keys2);
    w.idx = dp<uint32_t>(idx);
    w.idx_alt = dp<uint32_t>(idx2);
    w.count = dp<int64_t>(small);
    w.flag = reinterpret_cast<unsigned int *>(dp<int64_t>(small) + 2);
    w.tmp_bytes = sort_tmp_bytes(nt, end_bit, s);
    w.tmp = ctx->scratch_get(w.tmp_bytes);
    const int64_t nnz = sparse_phase_a(s, nt, I, J, m, n, end_bit, w);
    DevPtr start = dev_alloc(nnz * sizeof(int64_t));
    w.start = dp<int64_t>(start);
    M->nnz = nnz;
    M->rowval = dev_alloc(nnz * sizeof(int32_t));
    M->nzval = dev_alloc(nnz * sizeof(double));
    sparse_phase_b(s, nt, V, m, n, nnz, w, dp<int64_t>(M->colptr), dp<int32_t>(M->rowval), dp<double>(M->nzval));
    FEASM_HIP(hipStreamSynchronize(s));
    return M.release();
}

}  // namespace feasm
