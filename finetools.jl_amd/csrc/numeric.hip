// Numeric phase: global summation of the staged element blocks into the
// resident CSC values.
//
// The reference sums duplicates inside SparseArrays.sparse as a left fold in
// COO order (src/AssemblyModule.jl:319-325), i.e. for every global entry the
// element contributions are added in ascending element number.  Here one
// thread owns one node block (row node m, column node c) and walks its
// contribution list, which the symbolic phase stored in ascending element
// order: each of its ndn*ndn entries is therefore the same left fold, in the
// same order, as the reference's -- and it is computed without atomics, so
// the result is bitwise reproducible.  Accumulation is FP64 for both the FP64
// and the mixed-precision staging.
#include "common.hpp"

namespace feasm {

namespace {

template <int NDN, int SW, bool DOT, typename S>
__global__ __launch_bounds__(256) void k_gather(const int64_t *__restrict__ adj_ptr, const int32_t *__restrict__ bcol,
                                                const int64_t *__restrict__ cptr,
                                                const uint32_t *__restrict__ contrib, int64_t nblocks, int npairs,
                                                const S *__restrict__ stage, FormParams fp,
                                                double *__restrict__ nzval) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nblocks;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c = bcol[t];
        const int64_t b0 = adj_ptr[c];
        const int64_t deg = adj_ptr[c + 1] - b0;
        const int64_t r = t - b0;
        double acc[SW];
#pragma unroll
        for (int k = 0; k < SW; k++) acc[k] = 0.0;
        const int64_t i0 = cptr[t], i1 = cptr[t + 1];
        // contributions U at a time: their list entries, then their staged
        // blocks, are loaded before any is added, so the loads of one group
        // overlap; the sums still run in list order (the first value is
        // taken as it is, the left fold of sparse())
        constexpr int U = 4;
        for (int64_t i = i0; i < i1; i += U) {
            uint32_t u[U];
#pragma unroll
            for (int j = 0; j < U; j++) u[j] = i + j < i1 ? contrib[i + j] : 0u;
            S v[U][SW];
#pragma unroll
            for (int j = 0; j < U; j++) {
                const S *s = stage + ((int64_t)(u[j] >> 7) * npairs + ((u[j] >> 1) & 63u)) * SW;
                if (i + j < i1) {
#pragma unroll
                    for (int k = 0; k < SW; k++) v[j][k] = s[k];
                }
            }
#pragma unroll
            for (int j = 0; j < U; j++) {
                if (i + j >= i1) break;
                const bool tr = SW == 9 && (u[j] & 1u);
#pragma unroll
                for (int k = 0; k < SW; k++) {
                    // a transposed pair's block is read as its transpose
                    const double x = (double)(tr ? v[j][3 * (k % 3) + k / 3] : v[j][k]);
                    acc[k] = (i + j == i0) ? x : acc[k] + x;
                }
            }
        }
        // block (m, c): entry (k, l) sits in DOF column c*NDN + l at row slot r*NDN + k
        const int64_t base = (int64_t)NDN * NDN * b0 + r * NDN;
#pragma unroll
        for (int l = 0; l < NDN; l++)
#pragma unroll
            for (int k = 0; k < NDN; k++) {
                double v;
                if (DOT)
                    v = acc[0] * fp.coef[NDN * k + l];
                else
                    v = acc[SW == 9 ? 3 * k + l : 0];
                nzval[base + (int64_t)l * NDN * deg + k] = v;
            }
    }
}

template <int NDN, int SW, bool DOT, typename S>
void go(feasm_ctx *ctx, const feasm_mesh *m, const Pattern &p, const FormParams &fp, const void *stage,
        double *nzval) {
    hipLaunchKernelGGL((k_gather<NDN, SW, DOT, S>), dim3(grid_for(p.nblocks, 256)), dim3(256), 0, ctx->stream,
                       dp<int64_t>(p.adj_ptr), dp<int32_t>(p.bcol), dp<int64_t>(p.cptr), dp<uint32_t>(p.contrib),
                       p.nblocks, npairs(m->npe), static_cast<const S *>(stage), fp, nzval);
    FEASM_HIP(hipGetLastError());
}

template <typename S>
void by_form(feasm_ctx *ctx, const feasm_mesh *m, const Pattern &p, const FormParams &fp, const void *stage,
             double *nzval) {
    switch (fp.form) {
        case FEASM_FORM_LIN_ELASTIC: return go<3, 9, false, S>(ctx, m, p, fp, stage, nzval);
        case FEASM_FORM_DIFFUSION_ISO:
        case FEASM_FORM_DIFFUSION_MAT: return go<1, 1, false, S>(ctx, m, p, fp, stage, nzval);
        case FEASM_FORM_DOT:
            if (fp.ndn == 1) return go<1, 1, true, S>(ctx, m, p, fp, stage, nzval);
            if (fp.ndn == 2) return go<2, 1, true, S>(ctx, m, p, fp, stage, nzval);
            if (fp.ndn == 3) return go<3, 1, true, S>(ctx, m, p, fp, stage, nzval);
            break;
    }
    throw Error(FEASM_EUNSUPPORTED, "unsupported form / ndn in gather");
}

}  // namespace

void launch_gather(feasm_ctx *ctx, const feasm_mesh *m, const Pattern &p, const FormParams &fp, int precision,
                   const void *stage, double *nzval) {
    if (p.nblocks == 0) return;
    if (precision == FEASM_MIXED)
        by_form<float>(ctx, m, p, fp, stage, nzval);
    else
        by_form<double>(ctx, m, p, fp, stage, nzval);
}

}  // namespace feasm
