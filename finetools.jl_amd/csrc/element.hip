// Element integration kernels: the quadrature loops of the bilform_* element
// loops (src/FEMMBaseModule.jl:1335-1366, 1476-1535, 1774-1813) on gfx950.
//
// Mapping: one 64-lane wavefront per element (4 wavefronts per 256-thread
// workgroup, grid-stride over elements).  Per element:
//   A  lanes gather the npe x 3 node coordinates into LDS (gathervalues_asmat!,
//      src/FieldModule.jl:263-275);
//   B  lanes split the nqp x 9 Jacobian entries, J = X' * dN (locjac!,
//      src/MatrixUtilityModule.jl:60-68);
//   C  one lane per quadrature point forms det J by cofactors of row 1
//      (src/FESetModule.jl:479-489) and the adjugate inverse times 1/det
//      (gradN!, src/FESetModule.jl:507-528) -- always in FP64;
//   D  lanes split the nqp x npe spatial gradients (src/FESetModule.jl:537-543);
//   E  one lane per upper-triangle node pair (a <= b) integrates its block.
// Blocks are exactly symmetric by construction: a lane computes the (a, b)
// block once and the (b, a) block is its transpose, as complete_lt! makes the
// reference's element matrix (src/MatrixUtilityModule.jl:164-173).
//
// Forms (phase E):
//   elasticity  M = sum_q wJ g_a g_b^T, then K_ab[i][j] = sum_kl D[v(i,k)][v(j,l)] M[k][l],
//               which equals sum_q wJ B_a^T D B_b with B of blmat!
//               (src/DeforModelRedModule.jl:447-472) and strain order xx yy zz xy xz yz;
//   diffusion   S = sum_q (c J w) g_a . g_b          (add_mggt_ut_only!, :84-98)
//               S = sum_kl kappa[k][l] M[k][l]       (add_gkgt_ut_only!, :120-153)
//   dot         S = sum_q N_a N_b J w; the ndn x ndn coefficient is applied
//               where the blocks are summed (bilform_dot's factor * c[p,q]).
// Precision: FP64, or MIXED (gradients, block sums and staging in FP32; the
// Jacobian, its determinant and inverse stay FP64; the global summation is
// FP64 in the gather).
#include "common.hpp"

namespace feasm {

int npairs(int npe) { return npe * (npe + 1) / 2; }
int stage_width(int form) { return form == FEASM_FORM_LIN_ELASTIC ? 9 : 1; }

namespace {

constexpr int kWaves = 4;  // wavefronts per workgroup
constexpr int kBlock = 64 * kWaves;

// Orders one wavefront's LDS writes before its later reads: every wavefront
// works on its own LDS slices, so no workgroup barrier is needed.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Voigt index of the symmetric pair (i, k) in the strain order xx yy zz xy xz yz
__device__ constexpr int voigt(int i, int k) {
    return (i == k) ? i : ((i + k == 1) ? 3 : ((i + k == 2) ? 4 : 5));
}

// upper-triangle node pair p -> (a, b), a <= b, enumerated column by column:
// p = b(b+1)/2 + a
__device__ __forceinline__ void pair_of(int p, int &a, int &b) {
    int bb = (int)((sqrtf(8.0f * p + 1.0f) - 1.0f) * 0.5f);
    while ((bb + 1) * (bb + 2) / 2 <= p) bb++;
    while (bb * (bb + 1) / 2 > p) bb--;
    b = bb;
    a = p - bb * (bb + 1) / 2;
}

// Linear elasticity whose D has cubic symmetry (isotropic materials among
// them: D00 = D11 = D22, D01 = D02 = D12, D33 = D44 = D55, the rest zero):
// K = D01 M + D33 M^T off the diagonal, K_ii = D00 M_ii + D33 (tr M - M_ii).
constexpr int kElasticCubic = 16;

struct KParams {
    FormParams fp;
    const double *N;   // [nqp][npe]
    const double *dN;  // [nqp][npe][3]
    const double *w;   // [nqp]
};

enum OutMode { OUT_STAGE = 0, OUT_COO = 1 };

template <int NPE, int NQP, int FORM, typename R, int OUT>
__global__ __launch_bounds__(kBlock) void k_element(const int32_t *__restrict__ conn, const double *__restrict__ xyz,
                                                    int64_t nelem, KParams kp, R *__restrict__ stage,
                                                    const int32_t *__restrict__ dofs, int32_t *__restrict__ I,
                                                    int32_t *__restrict__ J, double *__restrict__ V) {
    constexpr int NP = NPE * (NPE + 1) / 2;
    constexpr bool GRAD = (FORM != FEASM_FORM_DOT);
    constexpr bool ELASTIC = (FORM == FEASM_FORM_LIN_ELASTIC || FORM == kElasticCubic);
    constexpr int SW = ELASTIC ? 9 : 1;
    __shared__ double sX[kWaves][NPE * 3];
    __shared__ double sJ[kWaves][NQP * 9];
    __shared__ R sInv[kWaves][NQP * 9];
    __shared__ R sW[kWaves][NQP];
    __shared__ R sG[kWaves][GRAD ? NQP * NPE * 3 : 1];

    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int ndn = kp.fp.ndn;
    const int kd = ndn * NPE;

    // Each wavefront walks its own elements (grid-stride) and synchronises
    // with itself only.
    for (int64_t e = (int64_t)blockIdx.x * kWaves + wv; e < nelem; e += (int64_t)gridDim.x * kWaves) {
        const int32_t *ce = conn + e * NPE;
        // A: coordinates
        for (int i = lane; i < NPE * 3; i += 64) sX[wv][i] = xyz[(int64_t)ce[i / 3] * 3 + (i % 3)];
        wave_sync();
        // B: Jacobian entries J[q][r][c] = sum_a X[a][r] dN[q][a][c]
        for (int i = lane; i < NQP * 9; i += 64) {
            const int q = i / 9, r = (i % 9) / 3, c = i % 3;
            const double *dq = kp.dN + q * NPE * 3;
            double s = 0.0;
#pragma unroll
            for (int a = 0; a < NPE; a++) s += sX[wv][3 * a + r] * dq[3 * a + c];
            sJ[wv][i] = s;
        }
        wave_sync();
        // C: determinant and inverse, FP64
        for (int q = lane; q < NQP; q += 64) {
            const double *Jq = &sJ[wv][q * 9];
#define RJ(i, j) Jq[3 * ((i)-1) + ((j)-1)]
            const double det = RJ(1, 1) * (RJ(2, 2) * RJ(3, 3) - RJ(3, 2) * RJ(2, 3)) -
                               RJ(1, 2) * (RJ(2, 1) * RJ(3, 3) - RJ(2, 3) * RJ(3, 1)) +
                               RJ(1, 3) * (RJ(2, 1) * RJ(3, 2) - RJ(2, 2) * RJ(3, 1));
            if (GRAD) {
                const double invdet = 1.0 / det;
                R *iv = &sInv[wv][q * 9];
                iv[0] = (R)((RJ(2, 2) * RJ(3, 3) - RJ(3, 2) * RJ(2, 3)) * invdet);
                iv[1] = (R)(-(RJ(1, 2) * RJ(3, 3) - RJ(1, 3) * RJ(3, 2)) * invdet);
                iv[2] = (R)((RJ(1, 2) * RJ(2, 3) - RJ(1, 3) * RJ(2, 2)) * invdet);
                iv[3] = (R)(-(RJ(2, 1) * RJ(3, 3) - RJ(2, 3) * RJ(3, 1)) * invdet);
                iv[4] = (R)((RJ(1, 1) * RJ(3, 3) - RJ(1, 3) * RJ(3, 1)) * invdet);
                iv[5] = (R)(-(RJ(1, 1) * RJ(2, 3) - RJ(2, 1) * RJ(1, 3)) * invdet);
                iv[6] = (R)((RJ(2, 1) * RJ(3, 2) - RJ(3, 1) * RJ(2, 2)) * invdet);
                iv[7] = (R)(-(RJ(1, 1) * RJ(3, 2) - RJ(3, 1) * RJ(1, 2)) * invdet);
                iv[8] = (R)((RJ(1, 1) * RJ(2, 2) - RJ(2, 1) * RJ(1, 2)) * invdet);
            }
#undef RJ
            double wj;
            if (FORM == FEASM_FORM_DIFFUSION_ISO)
                wj = (kp.fp.coef[0] * det) * kp.w[q];
            else
                wj = det * kp.w[q];
            sW[wv][q] = (R)wj;
        }
        wave_sync();
        // D: spatial gradients g[q][a][c] = sum_k dN[q][a][k] inv[k][c]
        if (GRAD) {
            for (int i = lane; i < NQP * NPE; i += 64) {
                const int q = i / NPE, a = i % NPE;
                const double *d = kp.dN + (q * NPE + a) * 3;
                const R d0 = (R)d[0], d1 = (R)d[1], d2 = (R)d[2];
                const R *iv = &sInv[wv][q * 9];
                R *g = &sG[wv][(q * NPE + a) * 3];
                g[0] = d0 * iv[0] + d1 * iv[3] + d2 * iv[6];
                g[1] = d0 * iv[1] + d1 * iv[4] + d2 * iv[7];
                g[2] = d0 * iv[2] + d1 * iv[5] + d2 * iv[8];
            }
            wave_sync();
        }
        // E: node-pair blocks
        for (int p = lane; p < NP; p += 64) {
            int a, b;
            pair_of(p, a, b);
            R out[SW];
            if (ELASTIC || FORM == FEASM_FORM_DIFFUSION_MAT) {
                R M[9];
#pragma unroll
                for (int i = 0; i < 9; i++) M[i] = (R)0;
#pragma unroll
                for (int q = 0; q < NQP; q++) {
                    const R wq = sW[wv][q];
                    const R *ga = &sG[wv][(q * NPE + a) * 3];
                    const R *gb = &sG[wv][(q * NPE + b) * 3];
#pragma unroll
                    for (int k = 0; k < 3; k++)
#pragma unroll
                        for (int l = 0; l < 3; l++) {
                            const R prod = ga[k] * gb[l];
                            M[3 * k + l] += wq * prod;
                        }
                }
                if (FORM == kElasticCubic) {
                    const R d00 = (R)kp.fp.coef[0], d01 = (R)kp.fp.coef[1], d33 = (R)kp.fp.coef[21];
                    const R tr = (M[0] + M[4]) + M[8];
#pragma unroll
                    for (int i = 0; i < 3; i++)
#pragma unroll
                        for (int j = 0; j < 3; j++)
                            out[3 * i + j] = (i == j) ? d00 * M[4 * i] + d33 * (tr - M[4 * i])
                                                      : d01 * M[3 * i + j] + d33 * M[3 * j + i];
                    if (a == b) {  // exact symmetry of the diagonal block
                        out[3] = out[1];
                        out[6] = out[2];
                        out[7] = out[5];
                    }
                } else if (FORM == FEASM_FORM_LIN_ELASTIC) {
#pragma unroll
                    for (int i = 0; i < 3; i++)
#pragma unroll
                        for (int j = 0; j < 3; j++) {
                            R s = (R)0;
#pragma unroll
                            for (int k = 0; k < 3; k++)
#pragma unroll
                                for (int l = 0; l < 3; l++)
                                    s += (R)kp.fp.coef[6 * voigt(i, k) + voigt(j, l)] * M[3 * k + l];
                            out[3 * i + j] = s;
                        }
                    if (a == b) {  // exact symmetry of the diagonal block
                        out[3] = out[1];
                        out[6] = out[2];
                        out[7] = out[5];
                    }
                } else {
                    R s = (R)0;
#pragma unroll
                    for (int k = 0; k < 3; k++)
#pragma unroll
                        for (int l = 0; l < 3; l++) s += (R)kp.fp.coef[3 * k + l] * M[3 * k + l];
                    out[0] = s;
                }
            } else if (FORM == FEASM_FORM_DIFFUSION_ISO) {
                R s = (R)0;
#pragma unroll
                for (int q = 0; q < NQP; q++) {
                    const R *ga = &sG[wv][(q * NPE + a) * 3];
                    const R *gb = &sG[wv][(q * NPE + b) * 3];
                    const R dot = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
                    s += sW[wv][q] * dot;
                }
                out[0] = s;
            } else {  // DOT
                R s = (R)0;
#pragma unroll
                for (int q = 0; q < NQP; q++) {
                    const R nn = (R)(kp.N[q * NPE + a] * kp.N[q * NPE + b]);
                    s += nn * sW[wv][q];
                }
                out[0] = s;
            }
            if (OUT == OUT_STAGE) {
                R *dst = stage + ((int64_t)e * NP + p) * SW;
#pragma unroll
                for (int s = 0; s < SW; s++) dst[s] = out[s];
            } else {
                // COO triples in assemble!'s column-major order
                // (src/AssemblyModule.jl:266-279): position j*kd + i.
                const int64_t base = e * (int64_t)kd * kd;
                const int na = ce[a], nb = ce[b];
                for (int k = 0; k < ndn; k++)
                    for (int l = 0; l < ndn; l++) {
                        double v;
                        if (ELASTIC)
                            v = (double)out[3 * k + l];
                        else if (FORM == FEASM_FORM_DOT)
                            v = (double)out[0] * kp.fp.coef[ndn * k + l];
                        else
                            v = (double)out[0];
                        const int ri = a * ndn + k, cj = b * ndn + l;
                        const int32_t dr = dofs[(int64_t)na * ndn + k], dc = dofs[(int64_t)nb * ndn + l];
                        int64_t pos = base + (int64_t)cj * kd + ri;
                        V[pos] = v;
                        I[pos] = dr;
                        J[pos] = dc;
                        if (a != b) {
                            // mirrored block: transpose for the symmetric forms,
                            // the same coefficient c[k][l] for the dot form
                            double vm = v;
                            int rk = b * ndn + l, ck = a * ndn + k;
                            int32_t mr = dc, mc = dr;
                            if (FORM == FEASM_FORM_DOT) {
                                rk = b * ndn + k;
                                ck = a * ndn + l;
                                mr = dofs[(int64_t)nb * ndn + k];
                                mc = dofs[(int64_t)na * ndn + l];
                            }
                            pos = base + (int64_t)ck * kd + rk;
                            V[pos] = vm;
                            I[pos] = mr;
                            J[pos] = mc;
                        }
                    }
            }
        }
        wave_sync();
    }
}

template <int NPE, int NQP, int FORM, typename R, int OUT>
void launch_t(feasm_ctx *ctx, const feasm_mesh *m, const KParams &kp, void *stage, const int32_t *dofs, int32_t *I,
              int32_t *J, double *V) {
    const int64_t groups = (m->nelem + kWaves - 1) / kWaves;
    const unsigned grid = grid_for(groups, 1, 1u << 18);
    hipLaunchKernelGGL((k_element<NPE, NQP, FORM, R, OUT>), dim3(grid), dim3(kBlock), 0, ctx->stream,
                       dp<int32_t>(m->conn), dp<double>(m->xyz), m->nelem, kp, static_cast<R *>(stage), dofs, I,
                       J, V);
    FEASM_HIP(hipGetLastError());
}

template <int NPE, int NQP, typename R, int OUT>
void by_form(feasm_ctx *ctx, const feasm_mesh *m, const KParams &kp, void *stage, const int32_t *dofs, int32_t *I,
             int32_t *J, double *V) {
    switch (kp.fp.form) {
        case FEASM_FORM_DIFFUSION_ISO:
            return launch_t<NPE, NQP, FEASM_FORM_DIFFUSION_ISO, R, OUT>(ctx, m, kp, stage, dofs, I, J, V);
        case FEASM_FORM_DIFFUSION_MAT:
            return launch_t<NPE, NQP, FEASM_FORM_DIFFUSION_MAT, R, OUT>(ctx, m, kp, stage, dofs, I, J, V);
        case FEASM_FORM_LIN_ELASTIC:
            if (cubic_symmetry(kp.fp.coef))
                return launch_t<NPE, NQP, kElasticCubic, R, OUT>(ctx, m, kp, stage, dofs, I, J, V);
            return launch_t<NPE, NQP, FEASM_FORM_LIN_ELASTIC, R, OUT>(ctx, m, kp, stage, dofs, I, J, V);
        case FEASM_FORM_DOT:
            return launch_t<NPE, NQP, FEASM_FORM_DOT, R, OUT>(ctx, m, kp, stage, dofs, I, J, V);
    }
    throw Error(FEASM_EINVAL, "unknown form");
}

template <typename R, int OUT>
void dispatch(feasm_ctx *ctx, const feasm_mesh *m, const DevQTable &qt, const FormParams &fp, void *stage,
              const int32_t *dofs, int32_t *I, int32_t *J, double *V) {
    KParams kp;
    kp.fp = fp;
    kp.N = qt.N();
    kp.dN = qt.dN();
    kp.w = qt.w();
    if (m->npe == 8) {
        switch (qt.nqp) {
            case 1: return by_form<8, 1, R, OUT>(ctx, m, kp, stage, dofs, I, J, V);
            case 8: return by_form<8, 8, R, OUT>(ctx, m, kp, stage, dofs, I, J, V);
            case 27: return by_form<8, 27, R, OUT>(ctx, m, kp, stage, dofs, I, J, V);
            case 64: return by_form<8, 64, R, OUT>(ctx, m, kp, stage, dofs, I, J, V);
        }
    } else if (m->npe == 10) {
        switch (qt.nqp) {
            case 1: return by_form<10, 1, R, OUT>(ctx, m, kp, stage, dofs, I, J, V);
            case 4: return by_form<10, 4, R, OUT>(ctx, m, kp, stage, dofs, I, J, V);
            case 5: return by_form<10, 5, R, OUT>(ctx, m, kp, stage, dofs, I, J, V);
        }
    }
    throw Error(FEASM_EUNSUPPORTED, "unsupported element/rule combination");
}

}  // namespace

void launch_element_stage(feasm_ctx *ctx, const feasm_mesh *m, const DevQTable &qt, const FormParams &fp,
                          int precision, void *stage) {
    if (precision == FEASM_MIXED)
        dispatch<float, OUT_STAGE>(ctx, m, qt, fp, stage, nullptr, nullptr, nullptr, nullptr);
    else
        dispatch<double, OUT_STAGE>(ctx, m, qt, fp, stage, nullptr, nullptr, nullptr, nullptr);
}

void launch_element_coo(feasm_ctx *ctx, const feasm_mesh *m, const DevQTable &qt, const FormParams &fp,
                        int precision, int32_t *I, int32_t *J, double *V) {
    if (precision == FEASM_MIXED)
        dispatch<float, OUT_COO>(ctx, m, qt, fp, nullptr, dp<int32_t>(m->dofs), I, J, V);
    else
        dispatch<double, OUT_COO>(ctx, m, qt, fp, nullptr, dp<int32_t>(m->dofs), I, J, V);
}

}  // namespace feasm
