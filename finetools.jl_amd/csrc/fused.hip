// Fused numeric assembly: element integration and the global summation in
// one kernel, nothing staged in HBM.
//
// One workgroup per tile (a compact set of owned column nodes plus the halo
// of elements incident to them; see tiles.cpp):
//   0  tile metadata, halo connectivity and the parametric gradient table
//      go to LDS;
//   1  one thread per (halo element, quadrature point): J = X' dN, det J by
//      cofactors of row 1 and the adjugate inverse (src/FESetModule.jl:479-544)
//      in FP64; the spatial gradients g_a = inv(J)' dN_a are stored scaled by
//      sqrt(|J w|) (u_a = sqrt(|Jw|) g_a, with the sign of J w kept apart), so
//      that an element's node-pair matrix is M_ab = sum_q s_q u_a u_b^T;
//   2  one thread per (owned column c, incident element e, half of the
//      element's nodes) forms the partial blocks M^e_{a c} (9 values each, or
//      one for the scalar forms) in registers and leaves them in LDS;
//   3  one thread per CSC node block (m, c) of the tile's columns adds the
//      partials of the elements holding both nodes in ascending element
//      number -- the order in which SparseArrays.sparse folds the duplicates
//      of the reference's element-major COO (src/AssemblyModule.jl:319-325) --
//      applies the material (K = D : M for elasticity,
//      src/DeforModelRedModule.jl:447-472 with src/MatrixUtilityModule.jl:
//      189-216) and writes the block once.
// The products u_a[k] u_c[l] are formed symmetrically (a sign flip is exact,
// the product commutes) and every block is contracted in the orientation of
// its lower node number, so the assembled matrix is exactly symmetric, as the
// reference's is (test/test_forms.jl:442).
#include "fused.hpp"

namespace feasm {

namespace {

constexpr int kMaxThreads = 512;

__device__ constexpr int voigt(int i, int k) {
    return (i == k) ? i : ((i + k == 1) ? 3 : ((i + k == 2) ? 4 : 5));
}

struct FusedParams {
    FormParams fp;
    const double *N;   // [nqp][npe]
    const double *dN;  // [nqp][npe][3]
    const double *w;   // [nqp]
};

// Everything one workgroup needs to process one tile of one mesh.
struct TileJob {
    const int32_t *node_ptr, *nodes, *node_cj, *halo_ptr, *halo, *cj_ptr, *hnode_ptr, *hnodes;
    const uint32_t *cj;
    const uint16_t *lconn;
    const int32_t *conn;
    const double *xyz;
    const int64_t *adj_ptr;
    const int32_t *adj;
    double *nzval;
    int64_t tile;
};

template <int NPE, int NQP, int FORM, int NDN, typename R>
__device__ __forceinline__ void fused_tile(const TileShape &shape, const TileJob &job, const FusedParams &kp,
                                           unsigned char *smem) {
    constexpr bool GRAD = (FORM != FEASM_FORM_DOT);
    constexpr int SWP = (FORM == FEASM_FORM_LIN_ELASTIC || FORM == FEASM_FORM_DIFFUSION_MAT) ? 9 : 1;
    constexpr int NSPLIT = 4;
    constexpr int APART = (NPE + NSPLIT - 1) / NSPLIT;
    const int32_t *__restrict__ node_ptr = job.node_ptr;
    const int32_t *__restrict__ hnodes = job.hnodes;
    const uint16_t *__restrict__ lconn = job.lconn;
    const int32_t *__restrict__ nodes = job.nodes;
    const int32_t *__restrict__ node_cj = job.node_cj;
    const int32_t *__restrict__ halo_ptr = job.halo_ptr;
    const int32_t *__restrict__ halo = job.halo;
    const int32_t *__restrict__ cj_ptr = job.cj_ptr;
    const uint32_t *__restrict__ cjg = job.cj;
    const int32_t *__restrict__ conn = job.conn;
    const double *__restrict__ xyz = job.xyz;
    const int64_t *__restrict__ adj_ptr = job.adj_ptr;
    const int32_t *__restrict__ adj = job.adj;
    double *__restrict__ nzval = job.nzval;
    const int64_t tile = job.tile;
    const int tid = threadIdx.x;
    const int kThreads = blockDim.x;

    const int n0 = node_ptr[tile], P = node_ptr[tile + 1] - n0;
    const int h0 = halo_ptr[tile], H = halo_ptr[tile + 1] - h0;
    const int c0 = cj_ptr[tile], CJ = cj_ptr[tile + 1] - c0;
    const int x0 = job.hnode_ptr[tile], NH = job.hnode_ptr[tile + 1] - x0;
    const TileShape::Layout L = shape.layout(H, CJ, P, NH);
    double *sDN = reinterpret_cast<double *>(smem + L.dN);
    double *sX = reinterpret_cast<double *>(smem + L.X);
    uint16_t *sLc = reinterpret_cast<uint16_t *>(smem + L.lconn);
    int32_t *sNode = reinterpret_cast<int32_t *>(smem + L.node);
    int32_t *sNcj = reinterpret_cast<int32_t *>(smem + L.ncj);
    int32_t *sBP = reinterpret_cast<int32_t *>(smem + L.bp);
    int32_t *sConn = reinterpret_cast<int32_t *>(smem + L.conn);
    R *sU = reinterpret_cast<R *>(smem + L.U);
    R *sS = reinterpret_cast<R *>(smem + L.S);
    uint32_t *sCJ = reinterpret_cast<uint32_t *>(smem + L.cj);
    R *sP = reinterpret_cast<R *>(smem + L.P);

    // ---- phase 0: metadata
    for (int i = tid; i < NQP * NPE * 3; i += kThreads) sDN[i] = kp.dN[i];
    for (int i = tid; i < P; i += kThreads) {
        sNode[i] = nodes[n0 + i];
        sNcj[i] = node_cj[n0 + i];
    }
    if (tid == 0) sNcj[P] = CJ;
    for (int i = tid; i < H * NPE; i += kThreads) {
        sConn[i] = conn[(int64_t)halo[h0 + i / NPE] * NPE + (i % NPE)];
        sLc[i] = lconn[(int64_t)h0 * NPE + i];
    }
    for (int i = tid; i < NH * 3; i += kThreads) sX[i] = xyz[(int64_t)hnodes[x0 + i / 3] * 3 + (i % 3)];
    for (int i = tid; i < CJ; i += kThreads) sCJ[i] = cjg[c0 + i];
    if (tid < 64) {  // block prefix over the owned columns (P <= 64): one wavefront scan
        int d = 0;
        if (tid < P) {
            const int32_t nd = nodes[n0 + tid];
            d = (int)(adj_ptr[nd + 1] - adj_ptr[nd]);
        }
        int s = d;
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(s, o, 64);
            if (tid >= o) s += v;
        }
        if (tid < P) sBP[tid + 1] = s;
        if (tid == 0) sBP[0] = 0;
    }
    __syncthreads();

    // ---- phase 1: Jacobians, inverses and scaled gradients
    for (int task = tid; task < H * NQP; task += kThreads) {
        const int h = task / NQP, q = task % NQP;
        const uint16_t *cn = sLc + h * NPE;
        const double *dq = sDN + q * NPE * 3;
        double J[9];
#pragma unroll
        for (int i = 0; i < 9; i++) J[i] = 0.0;
#pragma unroll
        for (int a = 0; a < NPE; a++) {
            const double *x = sX + (int)cn[a] * 3;
            const double x0 = x[0], x1 = x[1], x2 = x[2];
            const double d0 = dq[3 * a], d1 = dq[3 * a + 1], d2 = dq[3 * a + 2];
            J[0] += x0 * d0; J[1] += x0 * d1; J[2] += x0 * d2;
            J[3] += x1 * d0; J[4] += x1 * d1; J[5] += x1 * d2;
            J[6] += x2 * d0; J[7] += x2 * d1; J[8] += x2 * d2;
        }
#define RJ(i, j) J[3 * ((i)-1) + ((j)-1)]
        const double det = RJ(1, 1) * (RJ(2, 2) * RJ(3, 3) - RJ(3, 2) * RJ(2, 3)) -
                           RJ(1, 2) * (RJ(2, 1) * RJ(3, 3) - RJ(2, 3) * RJ(3, 1)) +
                           RJ(1, 3) * (RJ(2, 1) * RJ(3, 2) - RJ(2, 2) * RJ(3, 1));
        double wj = (FORM == FEASM_FORM_DIFFUSION_ISO) ? (kp.fp.coef[0] * det) * kp.w[q] : det * kp.w[q];
        if (GRAD) {
            const double invdet = 1.0 / det;
            const double iv[9] = {(RJ(2, 2) * RJ(3, 3) - RJ(3, 2) * RJ(2, 3)) * invdet,
                                  -(RJ(1, 2) * RJ(3, 3) - RJ(1, 3) * RJ(3, 2)) * invdet,
                                  (RJ(1, 2) * RJ(2, 3) - RJ(1, 3) * RJ(2, 2)) * invdet,
                                  -(RJ(2, 1) * RJ(3, 3) - RJ(2, 3) * RJ(3, 1)) * invdet,
                                  (RJ(1, 1) * RJ(3, 3) - RJ(1, 3) * RJ(3, 1)) * invdet,
                                  -(RJ(1, 1) * RJ(2, 3) - RJ(2, 1) * RJ(1, 3)) * invdet,
                                  (RJ(2, 1) * RJ(3, 2) - RJ(3, 1) * RJ(2, 2)) * invdet,
                                  -(RJ(1, 1) * RJ(3, 2) - RJ(3, 1) * RJ(1, 2)) * invdet,
                                  (RJ(1, 1) * RJ(2, 2) - RJ(2, 1) * RJ(1, 2)) * invdet};
            const R r = (R)sqrt(fabs(wj));
            sS[h * NQP + q] = (wj < 0.0) ? (R)-1 : (R)1;
            R ivr[9];
#pragma unroll
            for (int i = 0; i < 9; i++) ivr[i] = (R)iv[i];
            R *u = sU + (size_t)(h * NQP + q) * NPE * 3;
#pragma unroll
            for (int a = 0; a < NPE; a++) {
                const R d0 = (R)dq[3 * a], d1 = (R)dq[3 * a + 1], d2 = (R)dq[3 * a + 2];
                u[3 * a + 0] = r * (d0 * ivr[0] + d1 * ivr[3] + d2 * ivr[6]);
                u[3 * a + 1] = r * (d0 * ivr[1] + d1 * ivr[4] + d2 * ivr[7]);
                u[3 * a + 2] = r * (d0 * ivr[2] + d1 * ivr[5] + d2 * ivr[8]);
            }
        } else {
            sS[h * NQP + q] = (R)wj;
        }
#undef RJ
    }
    __syncthreads();

    // ---- phase 2: partial blocks of every (owned column, incident element)
    for (int task = tid; task < CJ * NSPLIT; task += kThreads) {
        const int cj = task / NSPLIT, part = task % NSPLIT;
        const uint32_t e = sCJ[cj];
        const int h = (int)((e >> 8) & 255u), ac = (int)(e & 255u);
        const int a0 = part * APART;
        R *out = sP + (size_t)cj * NPE * SWP;
        if (GRAD) {
            R uc[NQP][3], sq[NQP];
#pragma unroll
            for (int q = 0; q < NQP; q++) {
                const R *u = sU + ((size_t)(h * NQP + q) * NPE + ac) * 3;
                uc[q][0] = u[0];
                uc[q][1] = u[1];
                uc[q][2] = u[2];
                sq[q] = sS[h * NQP + q];
            }
#pragma unroll
            for (int ai = 0; ai < APART; ai++) {
                const int a = a0 + ai;
                if (a >= NPE) break;
                R M[SWP];
#pragma unroll
                for (int i = 0; i < SWP; i++) M[i] = (R)0;
#pragma unroll
                for (int q = 0; q < NQP; q++) {
                    const R *u = sU + ((size_t)(h * NQP + q) * NPE + a) * 3;
                    const R sa0 = sq[q] * u[0], sa1 = sq[q] * u[1], sa2 = sq[q] * u[2];
                    if (SWP == 9) {
                        M[0] = fma(sa0, uc[q][0], M[0]); M[1] = fma(sa0, uc[q][1], M[1]); M[2] = fma(sa0, uc[q][2], M[2]);
                        M[3] = fma(sa1, uc[q][0], M[3]); M[4] = fma(sa1, uc[q][1], M[4]); M[5] = fma(sa1, uc[q][2], M[5]);
                        M[6] = fma(sa2, uc[q][0], M[6]); M[7] = fma(sa2, uc[q][1], M[7]); M[8] = fma(sa2, uc[q][2], M[8]);
                    } else {
                        M[0] = fma(sa0, uc[q][0], M[0]);
                        M[0] = fma(sa1, uc[q][1], M[0]);
                        M[0] = fma(sa2, uc[q][2], M[0]);
                    }
                }
#pragma unroll
                for (int i = 0; i < SWP; i++) out[a * SWP + i] = M[i];
            }
        } else {
#pragma unroll
            for (int ai = 0; ai < APART; ai++) {
                const int a = a0 + ai;
                if (a >= NPE) break;
                R S = (R)0;
#pragma unroll
                for (int q = 0; q < NQP; q++) {
                    const R nn = (R)(kp.N[q * NPE + a] * kp.N[q * NPE + ac]);
                    S = fma(nn, sS[h * NQP + q], S);
                }
                out[a] = S;
            }
        }
    }
    __syncthreads();

    // ---- phase 3: ordered sums, material, one write per block
    const int NB = sBP[P];
    for (int task = tid; task < NB; task += kThreads) {
        int lo = 0, hi = P;  // column: last c with sBP[c] <= task
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (sBP[mid] <= task) lo = mid; else hi = mid;
        }
        const int c = lo;
        const int r = task - sBP[c];
        const int deg = sBP[c + 1] - sBP[c];
        const int32_t node = sNode[c];
        const int64_t b0 = adj_ptr[node];
        const int32_t m = adj[b0 + r];
        double acc[SWP];
#pragma unroll
        for (int i = 0; i < SWP; i++) acc[i] = 0.0;
        bool first = true;
        for (int cj = sNcj[c]; cj < sNcj[c + 1]; cj++) {
            const int h = (int)((sCJ[cj] >> 8) & 255u);
            const int32_t *cn = sConn + h * NPE;
            int am = -1;
#pragma unroll
            for (int a = 0; a < NPE; a++)
                if (cn[a] == m) am = a;
            if (am < 0) continue;
            const R *p = sP + ((size_t)cj * NPE + am) * SWP;
            if (first) {
#pragma unroll
                for (int i = 0; i < SWP; i++) acc[i] = (double)p[i];
                first = false;
            } else {
#pragma unroll
                for (int i = 0; i < SWP; i++) acc[i] = acc[i] + (double)p[i];
            }
        }
        double out[NDN * NDN];
        if (FORM == FEASM_FORM_LIN_ELASTIC || FORM == FEASM_FORM_DIFFUSION_MAT) {
            // contract in the orientation of the lower node number
            const bool sw = m > node;
            double Mc[9];
#pragma unroll
            for (int k = 0; k < 3; k++)
#pragma unroll
                for (int l = 0; l < 3; l++) Mc[3 * k + l] = sw ? acc[3 * l + k] : acc[3 * k + l];
            if (FORM == FEASM_FORM_LIN_ELASTIC) {
                double Kc[9];
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++) {
                        double s = 0.0;
#pragma unroll
                        for (int k = 0; k < 3; k++)
#pragma unroll
                            for (int l = 0; l < 3; l++) s += kp.fp.coef[6 * voigt(i, k) + voigt(j, l)] * Mc[3 * k + l];
                        Kc[3 * i + j] = s;
                    }
                if (m == node) {  // diagonal block: upper triangle mirrored, as complete_lt! does
                    Kc[3] = Kc[1];
                    Kc[6] = Kc[2];
                    Kc[7] = Kc[5];
                }
#pragma unroll
                for (int k = 0; k < 3; k++)
#pragma unroll
                    for (int l = 0; l < 3; l++) out[(NDN * k + l) % (NDN * NDN)] = sw ? Kc[3 * l + k] : Kc[3 * k + l];
            } else {
                double s = 0.0;
#pragma unroll
                for (int k = 0; k < 3; k++)
#pragma unroll
                    for (int l = 0; l < 3; l++) s += kp.fp.coef[3 * k + l] * Mc[3 * k + l];
                out[0] = s;
            }
        } else if (FORM == FEASM_FORM_DOT) {
#pragma unroll
            for (int k = 0; k < NDN; k++)
#pragma unroll
                for (int l = 0; l < NDN; l++) out[NDN * k + l] = acc[0] * kp.fp.coef[NDN * k + l];
        } else {
            out[0] = acc[0];
        }
        const int64_t base = (int64_t)NDN * NDN * b0 + (int64_t)r * NDN;
#pragma unroll
        for (int l = 0; l < NDN; l++)
#pragma unroll
            for (int k = 0; k < NDN; k++) nzval[base + (int64_t)l * NDN * deg + k] = out[NDN * k + l];
    }
    __syncthreads();  // LDS is reused by the workgroup's next tile
}

template <int NPE, int NQP, int FORM, int NDN, typename R>
__global__ __launch_bounds__(kMaxThreads) void k_fused(TileShape shape, TileJob job, int64_t ntiles, FusedParams kp) {
    extern __shared__ __align__(16) unsigned char smem[];
    // XCD-aware order: consecutive workgroups land on different XCDs, so hand
    // each XCD a contiguous range of (Morton-adjacent) tiles
    const int64_t per = (ntiles + 7) / 8;
    job.tile = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (job.tile >= ntiles) return;
    fused_tile<NPE, NQP, FORM, NDN, R>(shape, job, kp, smem);
}

// One entry per mesh of a batched launch.
struct BatchEntry {
    TileJob job;
    FusedParams kp;
    TileShape shape;
    int64_t tile0;
};

// Persistent batched launch: the workgroups of each XCD walk a contiguous
// range of the global tile list (all meshes' tiles, mesh after mesh); every
// tile carries its mesh's form, coefficients and arrays.
template <int NPE, int NQP, typename R>
__global__ __launch_bounds__(kMaxThreads) void k_fused_batch(const BatchEntry *__restrict__ entries,
                                                          const int32_t *__restrict__ tile_entry, int64_t ntiles) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int64_t per = (ntiles + 7) / 8;
    const int64_t xcd = blockIdx.x % 8, slot = blockIdx.x / 8, nslots = gridDim.x / 8;
    const int64_t begin = xcd * per;
    const int64_t end = (begin + per < ntiles) ? begin + per : ntiles;
    for (int64_t t = begin + slot; t < end; t += nslots) {
        const BatchEntry &E = entries[tile_entry[t]];
        TileJob job = E.job;
        job.tile = t - E.tile0;
        switch (E.kp.fp.form) {
            case FEASM_FORM_LIN_ELASTIC:
                fused_tile<NPE, NQP, FEASM_FORM_LIN_ELASTIC, 3, R>(E.shape, job, E.kp, smem);
                break;
            case FEASM_FORM_DIFFUSION_ISO:
                fused_tile<NPE, NQP, FEASM_FORM_DIFFUSION_ISO, 1, R>(E.shape, job, E.kp, smem);
                break;
            case FEASM_FORM_DIFFUSION_MAT:
                fused_tile<NPE, NQP, FEASM_FORM_DIFFUSION_MAT, 1, R>(E.shape, job, E.kp, smem);
                break;
            default:
                if (E.kp.fp.ndn == 3)
                    fused_tile<NPE, NQP, FEASM_FORM_DOT, 3, R>(E.shape, job, E.kp, smem);
                else if (E.kp.fp.ndn == 2)
                    fused_tile<NPE, NQP, FEASM_FORM_DOT, 2, R>(E.shape, job, E.kp, smem);
                else
                    fused_tile<NPE, NQP, FEASM_FORM_DOT, 1, R>(E.shape, job, E.kp, smem);
                break;
        }
    }
}

TileJob make_job(const feasm_mesh *m, const Pattern &p, const TileSet &t, double *nzval) {
    TileJob j;
    j.node_ptr = dp<int32_t>(t.node_ptr);
    j.nodes = dp<int32_t>(t.nodes);
    j.node_cj = dp<int32_t>(t.node_cj);
    j.halo_ptr = dp<int32_t>(t.halo_ptr);
    j.halo = dp<int32_t>(t.halo);
    j.cj_ptr = dp<int32_t>(t.cj_ptr);
    j.cj = dp<uint32_t>(t.cj);
    j.hnode_ptr = dp<int32_t>(t.hnode_ptr);
    j.hnodes = dp<int32_t>(t.hnodes);
    j.lconn = dp<uint16_t>(t.lconn);
    j.conn = dp<int32_t>(m->conn);
    j.xyz = dp<double>(m->xyz);
    j.adj_ptr = dp<int64_t>(p.adj_ptr);
    j.adj = dp<int32_t>(p.adj);
    j.nzval = nzval;
    j.tile = 0;
    return j;
}

template <int NPE, int NQP, int FORM, int NDN, typename R>
void go(feasm_ctx *ctx, const feasm_mesh *m, const Pattern &p, const TileSet &t, const FusedParams &kp,
        double *nzval) {
    auto kern = k_fused<NPE, NQP, FORM, NDN, R>;
    const int budget = (int)t.shape.budget;
    FEASM_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, budget));
    const int64_t per = (t.ntiles + 7) / 8;
    const unsigned grid = (unsigned)(8 * per);
    TileJob job = make_job(m, p, t, nzval);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(t.shape.threads), budget, ctx->stream, t.shape, job, t.ntiles, kp);
    FEASM_HIP(hipGetLastError());
}

template <int NPE, int NQP, typename R>
void by_form(feasm_ctx *ctx, const feasm_mesh *m, const Pattern &p, const TileSet &t, const FusedParams &kp,
             double *nzval) {
    switch (kp.fp.form) {
        case FEASM_FORM_LIN_ELASTIC: return go<NPE, NQP, FEASM_FORM_LIN_ELASTIC, 3, R>(ctx, m, p, t, kp, nzval);
        case FEASM_FORM_DIFFUSION_ISO: return go<NPE, NQP, FEASM_FORM_DIFFUSION_ISO, 1, R>(ctx, m, p, t, kp, nzval);
        case FEASM_FORM_DIFFUSION_MAT: return go<NPE, NQP, FEASM_FORM_DIFFUSION_MAT, 1, R>(ctx, m, p, t, kp, nzval);
        case FEASM_FORM_DOT:
            if (kp.fp.ndn == 1) return go<NPE, NQP, FEASM_FORM_DOT, 1, R>(ctx, m, p, t, kp, nzval);
            if (kp.fp.ndn == 2) return go<NPE, NQP, FEASM_FORM_DOT, 2, R>(ctx, m, p, t, kp, nzval);
            if (kp.fp.ndn == 3) return go<NPE, NQP, FEASM_FORM_DOT, 3, R>(ctx, m, p, t, kp, nzval);
            break;
    }
    throw Error(FEASM_EUNSUPPORTED, "unsupported form in the fused kernel");
}

template <typename R>
void dispatch(feasm_ctx *ctx, const feasm_mesh *m, const Pattern &p, const TileSet &t, const FusedParams &kp,
              double *nzval) {
    if (t.shape.npe == 8 && t.shape.nqp == 8) return by_form<8, 8, R>(ctx, m, p, t, kp, nzval);
    if (t.shape.npe == 10 && t.shape.nqp == 4) return by_form<10, 4, R>(ctx, m, p, t, kp, nzval);
    throw Error(FEASM_EUNSUPPORTED, "the fused kernel covers H8 with 2x2x2 Gauss and T10 with 4-point rules");
}

template <int NPE, int NQP, typename R>
void go_batch(feasm_ctx *ctx, const BatchLaunch &b) {
    auto kern = k_fused_batch<NPE, NQP, R>;
    FEASM_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b.budget));
    hipLaunchKernelGGL(kern, dim3(b.grid), dim3(b.threads), b.budget, ctx->stream,
                       static_cast<const BatchEntry *>(b.entries->p), dp<int32_t>(b.tile_entry), b.ntiles);
    FEASM_HIP(hipGetLastError());
}

}  // namespace

void prepare_batch_launch(feasm_ctx *ctx, BatchLaunch &b, const std::vector<BatchMesh> &items) {
    std::vector<BatchEntry> host;
    std::vector<int32_t> tile_entry;
    b.ntiles = 0;
    b.budget = 0;
    for (size_t i = 0; i < items.size(); i++) {
        const BatchMesh &it = items[i];
        BatchEntry E;
        E.job = make_job(it.mesh, *it.mesh->pattern, *it.tiles, it.nzval);
        E.kp.fp = it.fp;
        const DevQTable &qt = device_qtable(ctx, it.mesh->etype, it.rule, it.order);
        E.kp.N = qt.N();
        E.kp.dN = qt.dN();
        E.kp.w = qt.w();
        E.shape = it.tiles->shape;
        E.tile0 = b.ntiles;
        for (int64_t t = 0; t < it.tiles->ntiles; t++) tile_entry.push_back((int32_t)i);
        b.ntiles += it.tiles->ntiles;
        b.budget = std::max(b.budget, it.tiles->shape.budget);
        b.threads = it.tiles->shape.threads;
        host.push_back(E);
    }
    b.entries = dev_alloc(std::max<size_t>(host.size() * sizeof(BatchEntry), 16));
    b.tile_entry = dev_alloc(std::max<size_t>(tile_entry.size() * 4, 4));
    FEASM_HIP(hipMemcpy(b.entries->p, host.data(), host.size() * sizeof(BatchEntry), hipMemcpyHostToDevice));
    FEASM_HIP(hipMemcpy(b.tile_entry->p, tile_entry.data(), tile_entry.size() * 4, hipMemcpyHostToDevice));
    int cus = 0;
    FEASM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const size_t lds_cu = 160 * 1024;
    const int per_cu = (int)std::max<size_t>(1, lds_cu / std::max<size_t>(b.budget, 1));
    int64_t grid = (int64_t)cus * per_cu;
    grid = std::max<int64_t>(8, (grid / 8) * 8);
    const int64_t need = 8 * ((b.ntiles + 7) / 8);
    b.grid = (unsigned)std::min(grid, std::max<int64_t>(need, 8));
}

void launch_batch(feasm_ctx *ctx, const BatchLaunch &b) {
    if (b.ntiles == 0) return;
    if (b.npe == 8 && b.nqp == 8) {
        if (b.rbytes == 4) return go_batch<8, 8, float>(ctx, b);
        return go_batch<8, 8, double>(ctx, b);
    }
    if (b.npe == 10 && b.nqp == 4) {
        if (b.rbytes == 4) return go_batch<10, 4, float>(ctx, b);
        return go_batch<10, 4, double>(ctx, b);
    }
    throw Error(FEASM_EUNSUPPORTED, "batched launch covers H8 2x2x2 Gauss and T10 4-point rules");
}

bool fused_supported(int npe, int nqp) { return (npe == 8 && nqp == 8) || (npe == 10 && nqp == 4); }

TileShape tile_shape_for(feasm_ctx *ctx, int npe, int nqp, int form, int precision) {
    TileShape s;
    s.npe = npe;
    s.nqp = nqp;
    s.swp = (form == FEASM_FORM_LIN_ELASTIC || form == FEASM_FORM_DIFFUSION_MAT) ? 9 : 1;
    s.rbytes = (precision == FEASM_MIXED) ? 4 : 8;
    s.grad = (form != FEASM_FORM_DOT);
    int lds = 0;
    FEASM_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
    s.budget = (size_t)std::min(lds, 160 * 1024);
    s.max_nodes = 64;
    s.max_halo = 255;
    s.max_cj = 4096;
    s.threads = 512;
    return s;
}

void launch_fused(feasm_ctx *ctx, const feasm_mesh *m, const Pattern &p, const TileSet &t, const DevQTable &qt,
                  const FormParams &fp, int precision, double *nzval) {
    if (t.ntiles == 0) return;
    FusedParams kp;
    kp.fp = fp;
    kp.N = qt.N();
    kp.dN = qt.dN();
    kp.w = qt.w();
    if (precision == FEASM_MIXED)
        dispatch<float>(ctx, m, p, t, kp, nzval);
    else
        dispatch<double>(ctx, m, p, t, kp, nzval);
}

}  // namespace feasm
