// Load-vector assembly: linform_dot with a constant force on solid elements
// (m = 3) and the SysvecAssembler (src/FEMMBaseModule.jl:1207-1241,
// src/AssemblyModule.jl:853-917).
//
// k_vec_elem: one thread per element forms the element vector exactly as the
// reference loop does -- per quadrature point Factor = det J * w, NkxF = N_k *
// Factor, elvec[(k, mx)] += NkxF * force[mx] -- and stages it in HBM.
// (k_vec_factor / k_vec_factor_gather below do the same with less staging
// for H8 and T10 with their usual rules.)
// k_vec_gather: one thread per node adds, for each of its DOFs, the staged
// entries of its elements in ascending element number (the node -> element
// lists are sorted), starting from zero: the same additions, in the same
// order, as assemble!'s F[gi] += vec[i] over the element loop.
#include "common.hpp"

namespace feasm {

namespace {

template <int NPE>
__global__ __launch_bounds__(256) void k_vec_elem(const int32_t *__restrict__ conn, const double *__restrict__ xyz,
                                                  int64_t nelem, int nqp, const double *__restrict__ N,
                                                  const double *__restrict__ dN, const double *__restrict__ w,
                                                  double f0, double f1, double f2, int ndn,
                                                  double *__restrict__ stage) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nelem; e += (int64_t)gridDim.x * blockDim.x) {
    double X[NPE][3];
#pragma unroll
    for (int a = 0; a < NPE; a++) {
        const double *p = xyz + 3 * (int64_t)conn[e * NPE + a];
        X[a][0] = p[0];
        X[a][1] = p[1];
        X[a][2] = p[2];
    }
    const double force[3] = {f0, f1, f2};
    double ev[NPE * 3];
#pragma unroll
    for (int i = 0; i < NPE * 3; i++) ev[i] = 0.0;
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
        const double det = J[0] * (J[4] * J[8] - J[7] * J[5]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                           J[2] * (J[3] * J[7] - J[4] * J[6]);
        const double Factor = det * w[q];
#pragma unroll
        for (int a = 0; a < NPE; a++) {
            const double NkxF = N[q * NPE + a] * Factor;
#pragma unroll
            for (int k = 0; k < 3; k++)
                if (k < ndn) ev[a * 3 + k] = ev[a * 3 + k] + NkxF * force[k];
        }
    }
    double *out = stage + e * NPE * ndn;
#pragma unroll
    for (int a = 0; a < NPE; a++)
#pragma unroll
        for (int k = 0; k < 3; k++)
            if (k < ndn) out[a * ndn + k] = ev[a * 3 + k];
  }
}

__global__ __launch_bounds__(256) void k_vec_gather(const int64_t *__restrict__ n2e_ptr,
                                                    const uint32_t *__restrict__ n2e, int64_t nnodes, int npe,
                                                    int ndn, const double *__restrict__ stage,
                                                    const int32_t *__restrict__ dofs, double *__restrict__ F) {
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < nnodes; n += (int64_t)gridDim.x * blockDim.x) {
    if (n2e_ptr[n] == n2e_ptr[n + 1]) continue;  // a node of no element: its DOFs are never touched
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t s = n2e_ptr[n]; s < n2e_ptr[n + 1]; s++) {
        const uint32_t v = n2e[s];
        const int64_t e = v >> 4;
        const int a = (int)(v & 15u);
        const double *p = stage + (e * npe + a) * ndn;
        for (int k = 0; k < ndn; k++) acc[k] = acc[k] + p[k];
    }
    for (int k = 0; k < ndn; k++) F[dofs[n * ndn + k]] = acc[k];
  }
}

// The common element families (H8 with 8 points, T10 with 4): the element
// kernel stages only Factor = det J * w per point, and the gather forms each
// element vector entry from it -- NkxF = N_k * Factor, elvec += NkxF *
// force[mx] over the points in order -- before adding it to the node's sum:
// the same operations, in the same order, as the staged element vectors.
template <int NPE, int NQP>
__global__ __launch_bounds__(256) void k_vec_factor(const int32_t *__restrict__ conn, const double *__restrict__ xyz,
                                                    int64_t nelem, const double *__restrict__ dN,
                                                    const double *__restrict__ w, double *__restrict__ fac) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nelem;
         e += (int64_t)gridDim.x * blockDim.x) {
        double X[NPE][3];
#pragma unroll
        for (int a = 0; a < NPE; a++) {
            const double *p = xyz + 3 * (int64_t)conn[e * NPE + a];
            X[a][0] = p[0];
            X[a][1] = p[1];
            X[a][2] = p[2];
        }
#pragma unroll
        for (int q = 0; q < NQP; q++) {
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
            const double det = J[0] * (J[4] * J[8] - J[7] * J[5]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                               J[2] * (J[3] * J[7] - J[4] * J[6]);
            fac[e * NQP + q] = det * w[q];
        }
    }
}

template <int NPE, int NQP>
__global__ __launch_bounds__(256) void k_vec_factor_gather(const int64_t *__restrict__ n2e_ptr,
                                                           const uint32_t *__restrict__ n2e, int64_t nnodes,
                                                           const double *__restrict__ N, int ndn, double f0,
                                                           double f1, double f2, const double *__restrict__ fac,
                                                           const int32_t *__restrict__ dofs, double *__restrict__ F) {
    __shared__ double sN[NQP * NPE];
    for (int i = threadIdx.x; i < NQP * NPE; i += blockDim.x) sN[i] = N[i];
    __syncthreads();
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < nnodes;
         n += (int64_t)gridDim.x * blockDim.x) {
        if (n2e_ptr[n] == n2e_ptr[n + 1]) continue;  // a node of no element: its DOFs are never touched
        double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
        // the node's elements four at a time: their list entries, then their
        // point factors, are loaded before any is used (the sums still run
        // in list order)
        constexpr int U = 4;
        const int64_t s0 = n2e_ptr[n], s1 = n2e_ptr[n + 1];
        for (int64_t s = s0; s < s1; s += U) {
            uint32_t v[U];
#pragma unroll
            for (int j = 0; j < U; j++) v[j] = s + j < s1 ? n2e[s + j] : 0u;
            double fq[U][NQP];
#pragma unroll
            for (int j = 0; j < U; j++) {
                const double *fe = fac + (int64_t)(v[j] >> 4) * NQP;
                if (s + j < s1) {
#pragma unroll
                    for (int q = 0; q < NQP; q++) fq[j][q] = fe[q];
                }
            }
#pragma unroll
            for (int j = 0; j < U; j++) {
                if (s + j >= s1) break;
                const int a = (int)(v[j] & 15u);
                double ev0 = 0.0, ev1 = 0.0, ev2 = 0.0;
#pragma unroll
                for (int q = 0; q < NQP; q++) {
                    const double NkxF = sN[q * NPE + a] * fq[j][q];
                    ev0 = ev0 + NkxF * f0;
                    ev1 = ev1 + NkxF * f1;
                    ev2 = ev2 + NkxF * f2;
                }
                acc0 = acc0 + ev0;
                acc1 = acc1 + ev1;
                acc2 = acc2 + ev2;
            }
        }
        const int32_t *d = dofs + n * ndn;
        F[d[0]] = acc0;
        if (ndn > 1) F[d[1]] = acc1;
        if (ndn > 2) F[d[2]] = acc2;
    }
}

}  // namespace

void launch_linform_dot(feasm_ctx *ctx, const feasm_mesh *m, const DevQTable &qt, const double *force,
                        double *stage, double *F) {
    FEASM_REQUIRE(m->conn && m->xyz && m->n2e_ptr && m->n2e && m->dofs && stage && F, FEASM_ESTATE,
                  "load vector: mesh arrays missing");
    const int ndn = m->ndn;
    const double f0 = force[0], f1 = ndn > 1 ? force[1] : 0.0, f2 = ndn > 2 ? force[2] : 0.0;
    FEASM_HIP(hipMemsetAsync(F, 0, (size_t)m->nalldofs * sizeof(double), ctx->stream));
    const unsigned ge = grid_for(m->nelem, 256);
    if ((m->npe == 8 && qt.nqp == 8) || (m->npe == 10 && qt.nqp == 4)) {  // stage holds nelem*npe*ndn >= nelem*nqp
        const unsigned gn = grid_for(m->nnodes, 256);
        if (m->npe == 8) {
            hipLaunchKernelGGL((k_vec_factor<8, 8>), dim3(ge), dim3(256), 0, ctx->stream, dp<int32_t>(m->conn),
                               dp<double>(m->xyz), m->nelem, qt.dN(), qt.w(), stage);
            FEASM_HIP(hipGetLastError());
            hipLaunchKernelGGL((k_vec_factor_gather<8, 8>), dim3(gn), dim3(256), 0, ctx->stream,
                               dp<int64_t>(m->n2e_ptr), dp<uint32_t>(m->n2e), m->nnodes, qt.N(), ndn, f0, f1, f2,
                               (const double *)stage, dp<int32_t>(m->dofs), F);
        } else {
            hipLaunchKernelGGL((k_vec_factor<10, 4>), dim3(ge), dim3(256), 0, ctx->stream, dp<int32_t>(m->conn),
                               dp<double>(m->xyz), m->nelem, qt.dN(), qt.w(), stage);
            FEASM_HIP(hipGetLastError());
            hipLaunchKernelGGL((k_vec_factor_gather<10, 4>), dim3(gn), dim3(256), 0, ctx->stream,
                               dp<int64_t>(m->n2e_ptr), dp<uint32_t>(m->n2e), m->nnodes, qt.N(), ndn, f0, f1, f2,
                               (const double *)stage, dp<int32_t>(m->dofs), F);
        }
        FEASM_HIP(hipGetLastError());
        return;
    }
    if (m->npe == 8)
        hipLaunchKernelGGL((k_vec_elem<8>), dim3(ge), dim3(256), 0, ctx->stream, dp<int32_t>(m->conn),
                           dp<double>(m->xyz), m->nelem, qt.nqp, qt.N(), qt.dN(), qt.w(), f0, f1, f2, ndn, stage);
    else if (m->npe == 10)
        hipLaunchKernelGGL((k_vec_elem<10>), dim3(ge), dim3(256), 0, ctx->stream, dp<int32_t>(m->conn),
                           dp<double>(m->xyz), m->nelem, qt.nqp, qt.N(), qt.dN(), qt.w(), f0, f1, f2, ndn, stage);
    else
        throw Error(FEASM_EUNSUPPORTED, "load vectors cover H8 and T10 meshes");
    FEASM_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_vec_gather, dim3(grid_for(m->nnodes, 256)), dim3(256), 0, ctx->stream,
                       dp<int64_t>(m->n2e_ptr), dp<uint32_t>(m->n2e), m->nnodes, m->npe, ndn, stage,
                       dp<int32_t>(m->dofs), F);
    FEASM_HIP(hipGetLastError());
}

}  // namespace feasm
