// HRZ-lumped mass: bilform_dot (src/FEMMBaseModule.jl:1335-1366) into a
// SysmatAssemblerSparseHRZLumpingSymm (src/AssemblyModule.jl:943-1141).
//
// k_hrz_elem: one thread per element forms the element mass matrix entries in
// the reference's arithmetic -- per quadrature point factor = N_a N_b J w,
// Ke += factor * c -- only as far as the lumping needs them: the column sums
// summed (em2), the diagonal summed (dem2), and the diagonal scaled by
// em2 / dem2, which it stages per (element, node, component).  The DOFs per
// node (1 to 3) are a template argument, so the node pair's ndn x ndn entries
// live in registers.
// k_hrz_gather: one thread per node adds its elements' staged values in
// ascending element number (sparse() meets the element-major triples in that
// order) and marks the DOFs that received a value.
// k_hrz_compact: the diagonal CSC of the marked DOFs.
#include "common.hpp"

namespace feasm {

namespace {

// det J of every quadrature point goes to the thread's column of the block's
// LDS (jac[q * blockDim + thread]); the element matrix is then visited column
// node by column node (b), each over the points (q) and a group of row nodes
// (a) at once, so that each point's factor N_a N_b J w is formed once for all
// ndn x ndn components of the node pair.  Every entry still sums its points in
// ascending order and the column sums add rows in (a, p) order, as the
// reference's Ke += factor * c loop and sum(Ke) give them.
template <int NPE, int NDN>
__global__ __launch_bounds__(256) void k_hrz_elem(const int32_t *__restrict__ conn, const double *__restrict__ xyz,
                                                  int64_t nelem, int nqp, const double *__restrict__ N,
                                                  const double *__restrict__ dN, const double *__restrict__ w,
                                                  const double *__restrict__ coef, double *__restrict__ stage) {
  extern __shared__ double jac[];
  const int nt = blockDim.x;
  double *const jt = jac + threadIdx.x;
  double c[NDN][NDN];
#pragma unroll
  for (int p = 0; p < NDN; p++)
#pragma unroll
      for (int k = 0; k < NDN; k++) c[p][k] = coef[NDN * p + k];
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nelem; e += (int64_t)gridDim.x * blockDim.x) {
    {
        double X[NPE][3];
#pragma unroll
        for (int a = 0; a < NPE; a++) {
            const double *p = xyz + 3 * (int64_t)conn[e * NPE + a];
            X[a][0] = p[0];
            X[a][1] = p[1];
            X[a][2] = p[2];
        }
        for (int q = 0; q < nqp; q++) {
            double J[9];
#pragma unroll
            for (int i = 0; i < 9; i++) J[i] = 0.0;
#pragma unroll
            for (int a = 0; a < NPE; a++) {
                const double d0 = dN[(q * NPE + a) * 3], d1 = dN[(q * NPE + a) * 3 + 1], d2 = dN[(q * NPE + a) * 3 + 2];
                J[0] += X[a][0] * d0; J[1] += X[a][0] * d1; J[2] += X[a][0] * d2;
                J[3] += X[a][1] * d0; J[4] += X[a][1] * d1; J[5] += X[a][1] * d2;
                J[6] += X[a][2] * d0; J[7] += X[a][2] * d1; J[8] += X[a][2] * d2;
            }
            jt[q * nt] = J[0] * (J[4] * J[8] - J[7] * J[5]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                         J[2] * (J[3] * J[7] - J[4] * J[6]);
        }
    }
    // em2: the column (b, cq) sums of Ke[(a, p), (b, cq)] = sum_q (N_a N_b J w) c[p][cq], added up
    double em2 = 0.0;
    // the row nodes in groups of AC (fewer live accumulators); the column
    // sums still add the rows in ascending (a, p) order
    constexpr int AC = NDN == 3 ? NPE / 2 : NPE;
    for (int b = 0; b < NPE; b++) {
        double cs[NDN];
#pragma unroll
        for (int k = 0; k < NDN; k++) cs[k] = 0.0;
#pragma unroll
        for (int a0 = 0; a0 < NPE; a0 += AC) {
            double ke[AC][NDN][NDN];
#pragma unroll
            for (int i = 0; i < AC; i++)
#pragma unroll
                for (int p = 0; p < NDN; p++)
#pragma unroll
                    for (int k = 0; k < NDN; k++) ke[i][p][k] = 0.0;
            for (int q = 0; q < nqp; q++) {
                const double nb = N[q * NPE + b], jq = jt[q * nt], wq = w[q];
#pragma unroll
                for (int i = 0; i < AC; i++) {
                    const double factor = N[q * NPE + a0 + i] * nb * jq * wq;
#pragma unroll
                    for (int p = 0; p < NDN; p++)
#pragma unroll
                        for (int k = 0; k < NDN; k++) ke[i][p][k] += factor * c[p][k];
                }
            }
#pragma unroll
            for (int i = 0; i < AC; i++) {
#pragma unroll
                for (int p = 0; p < NDN; p++)
#pragma unroll
                    for (int k = 0; k < NDN; k++) cs[k] += ke[i][p][k];
            }
        }
#pragma unroll
        for (int k = 0; k < NDN; k++) em2 += cs[k];
    }
    // the diagonal entries Ke[(a, p), (a, p)], formed as the loop above
    // forms them for b = a (the same operations, so the same values)
    double diag[NPE][NDN];
#pragma unroll
    for (int a = 0; a < NPE; a++)
#pragma unroll
        for (int p = 0; p < NDN; p++) diag[a][p] = 0.0;
    for (int q = 0; q < nqp; q++) {
        const double jq = jt[q * nt], wq = w[q];
#pragma unroll
        for (int a = 0; a < NPE; a++) {
            const double na = N[q * NPE + a];
            const double factor = na * na * jq * wq;
#pragma unroll
            for (int p = 0; p < NDN; p++) diag[a][p] += factor * c[p][p];
        }
    }
    double dem2 = 0.0;
#pragma unroll
    for (int a = 0; a < NPE; a++)
#pragma unroll
        for (int p = 0; p < NDN; p++) dem2 += diag[a][p];
    const double ffactor = em2 / dem2;
    double *out = stage + e * NPE * NDN;
#pragma unroll
    for (int a = 0; a < NPE; a++)
#pragma unroll
        for (int p = 0; p < NDN; p++) out[a * NDN + p] = diag[a][p] * ffactor;
  }
}

// The node's list entries are read eight at a time (all of an interior H8
// node's), so the loads of a group are in flight together; the values are
// still added in list order.
template <int NDN>
__global__ __launch_bounds__(256) void k_hrz_gather(const int64_t *__restrict__ n2e_ptr,
                                                    const uint32_t *__restrict__ n2e, int64_t nnodes, int npe,
                                                    const double *__restrict__ stage,
                                                    const int32_t *__restrict__ dofs, double *__restrict__ val,
                                                    int32_t *__restrict__ flag) {
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < nnodes; n += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s0 = n2e_ptr[n], s1 = n2e_ptr[n + 1];
    if (s0 == s1) continue;  // a node of no element: no triple, no stored entry
    double acc[NDN];
#pragma unroll
    for (int k = 0; k < NDN; k++) acc[k] = 0.0;
    constexpr int U = 8;
    for (int64_t s = s0; s < s1; s += U) {
        uint32_t v[U];
#pragma unroll
        for (int j = 0; j < U; j++) v[j] = s + j < s1 ? n2e[s + j] : 0u;
        double x[U][NDN];
#pragma unroll
        for (int j = 0; j < U; j++) {
            const double *p = stage + ((int64_t)(v[j] >> 4) * npe + (v[j] & 15u)) * NDN;
            if (s + j < s1) {
#pragma unroll
                for (int k = 0; k < NDN; k++) x[j][k] = p[k];
            }
        }
#pragma unroll
        for (int j = 0; j < U; j++) {
            if (s + j >= s1) break;
#pragma unroll
            for (int k = 0; k < NDN; k++) acc[k] = acc[k] + x[j][k];
        }
    }
#pragma unroll
    for (int k = 0; k < NDN; k++) {
        const int32_t d = dofs[n * NDN + k];
        val[d] = acc[k];
        flag[d] = 1;
    }
  }
}

__global__ __launch_bounds__(256) void k_hrz_compact(const int32_t *__restrict__ flag, const int64_t *__restrict__ pos,
                                                     const double *__restrict__ val, int64_t n,
                                                     int32_t *__restrict__ rowval, double *__restrict__ nzval) {
    for (int64_t d = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; d < n; d += (int64_t)gridDim.x * blockDim.x)
        if (flag[d]) {
            rowval[pos[d]] = (int32_t)d;
            nzval[pos[d]] = val[d];
        }
}

}  // namespace

feasm_matrix *hrz_lumped(feasm_ctx *ctx, feasm_mesh *m, const DevQTable &qt, const double *coef_host) {
    const int ndn = m->ndn;
    const int64_t n = m->nalldofs;
    FEASM_REQUIRE(m->conn && m->xyz && m->n2e_ptr && m->n2e && m->dofs, FEASM_ESTATE, "lumped mass: mesh arrays missing");
    FEASM_REQUIRE(qt.nqp <= 64, FEASM_EUNSUPPORTED, "lumped mass: rules of at most 64 points");
    FEASM_REQUIRE(ndn >= 1 && ndn <= 3, FEASM_EUNSUPPORTED, "lumped mass: 1 to 3 DOFs per node");
    DevPtr coef = dev_alloc(ndn * ndn * sizeof(double));
    FEASM_HIP(hipMemcpyAsync(coef->p, coef_host, ndn * ndn * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    DevPtr stage = dev_alloc(std::max<size_t>((size_t)m->nelem * m->npe * ndn * sizeof(double), 8));
    DevPtr val = dev_alloc(std::max<int64_t>(n, 1) * sizeof(double));
    DevPtr flag = dev_alloc((n + 1) * sizeof(int32_t));
    FEASM_HIP(hipMemsetAsync(flag->p, 0, (n + 1) * sizeof(int32_t), ctx->stream));
    // det J per (point, thread) in LDS: at most 32 KB per block
    const int be = qt.nqp <= 16 ? 256 : 64;
    const unsigned ge = grid_for(m->nelem, be);
    const size_t lds = (size_t)qt.nqp * be * sizeof(double);
    auto launch = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(ge), dim3(be), lds, ctx->stream, dp<int32_t>(m->conn), dp<double>(m->xyz),
                           m->nelem, qt.nqp, qt.N(), qt.dN(), qt.w(), dp<double>(coef), dp<double>(stage));
    };
    if (m->npe == 8)
        ndn == 1 ? launch(k_hrz_elem<8, 1>) : ndn == 2 ? launch(k_hrz_elem<8, 2>) : launch(k_hrz_elem<8, 3>);
    else if (m->npe == 10)
        ndn == 1 ? launch(k_hrz_elem<10, 1>) : ndn == 2 ? launch(k_hrz_elem<10, 2>) : launch(k_hrz_elem<10, 3>);
    else
        throw Error(FEASM_EUNSUPPORTED, "lumped mass covers H8 and T10 meshes");
    FEASM_HIP(hipGetLastError());
    auto gather = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(grid_for(m->nnodes, 256)), dim3(256), 0, ctx->stream, dp<int64_t>(m->n2e_ptr),
                           dp<uint32_t>(m->n2e), m->nnodes, m->npe, dp<double>(stage), dp<int32_t>(m->dofs),
                           dp<double>(val), dp<int32_t>(flag));
    };
    ndn == 1 ? gather(k_hrz_gather<1>) : ndn == 2 ? gather(k_hrz_gather<2>) : gather(k_hrz_gather<3>);
    FEASM_HIP(hipGetLastError());
    std::unique_ptr<feasm_matrix> M(new feasm_matrix);
    M->ctx = ctx;
    M->m = M->n = n;
    M->colptr = dev_alloc((n + 1) * sizeof(int64_t));
    exclusive_scan_i32_to_i64(ctx, dp<int32_t>(flag), dp<int64_t>(M->colptr), n + 1);
    int64_t nnz = 0;
    read_back(ctx->stream, &nnz, dp<int64_t>(M->colptr) + n, sizeof(int64_t));
    M->nnz = nnz;
    M->rowval = dev_alloc(std::max<int64_t>(nnz, 1) * sizeof(int32_t));
    M->nzval = dev_alloc(std::max<int64_t>(nnz, 1) * sizeof(double));
    hipLaunchKernelGGL(k_hrz_compact, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, dp<int32_t>(flag),
                       dp<int64_t>(M->colptr), dp<double>(val), n, dp<int32_t>(M->rowval), dp<double>(M->nzval));
    FEASM_HIP(hipGetLastError());
    FEASM_HIP(hipStreamSynchronize(ctx->stream));
    return M.release();
}

}  // namespace feasm
