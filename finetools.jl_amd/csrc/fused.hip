// Fused numeric assembly: element integration and the global summation in
// one kernel, nothing staged in HBM.
//
// Each workgroup walks a contiguous range of tiles (a compact set of owned
// column nodes plus the halo of elements incident to them, with the tile's
// node blocks listed as tasks; see tiles.cpp).  Per tile:
//   0  the tile's LDS image (column pointers, task and item lists, halo
//      connectivity; one contiguous blob laid out by TileShape::layout) and
//      the coordinates of its halo nodes go to LDS.  These global loads are
//      issued while the previous tile is in phase 2 and held in registers;
//   1  one thread per (halo element, quadrature point): J = X' dN, det J by
//      cofactors of row 1 and the adjugate inverse (src/FESetModule.jl:479-544)
//      in FP64; the spatial gradients g_a = inv(J)' dN_a are stored scaled by
//      sqrt(|J w|) (u_a = sqrt(|Jw|) g_a, with the sign of J w kept apart), so
//      that an element's node-pair matrix is M_ab = sum_q s_q u_a u_b^T;
//   2  one thread per CSC node block (m, c) of the tile's columns forms the
//      element matrices M^e_{m c} of the elements holding both nodes and adds
//      them in ascending element number -- the order in which
//      SparseArrays.sparse folds the duplicates of the reference's
//      element-major COO (src/AssemblyModule.jl:319-325) -- applies the
//      material (K = D : M for elasticity, src/DeforModelRedModule.jl:447-472
//      with src/MatrixUtilityModule.jl:189-216) and writes the block once.
// The products u_m[k] u_c[l] are formed symmetrically (a sign flip is exact,
// the product commutes) and every block is contracted in the orientation of
// its lower node number, so the assembled matrix is exactly symmetric, as the
// reference's is (test/test_forms.jl:442).
#include "fused.hpp"

namespace feasm {

namespace {

constexpr int kMaxThreads = 1024;

__device__ constexpr int voigt(int i, int k) {
    return (i == k) ? i : ((i + k == 1) ? 3 : ((i + k == 2) ? 4 : 5));
}

// Linear elasticity with a material matrix of cubic symmetry (isotropic
// materials among them): D00 = D11 = D22, D01 = D02 = D12, D33 = D44 = D55,
// every other entry zero.  Then K = D01 M + D33 M^T off the diagonal and
// K_ii = D00 M_ii + D33 (tr M - M_ii), three coefficients instead of 36.
constexpr int kFormElasticCubic = 16;

}  // namespace

bool cubic_symmetry(const double *D) {
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            const double v = D[6 * i + j];
            if (i < 3 && j < 3) {
                if (v != (i == j ? D[0] : D[1])) return false;
            } else if (i == j) {
                if (v != D[21]) return false;
            } else if (v != 0.0) {
                return false;
            }
        }
    return true;
}

namespace {

struct FusedParams {
    FormParams fp;
    int cubic;         // elasticity with a cubic-symmetry D (see kFormElasticCubic)
    const double *N;   // [nqp][npe]
    const double *dN;  // [nqp][npe][3]
    const double *w;   // [nqp]
};

// Everything one workgroup needs to process the tiles of one mesh.
struct TileJob {
    const TileHdr *hdr;
    const unsigned char *blob;
    const int32_t *hnodes;
    const double *xyz;
    double *nzval;
};

// signs of J w at the points of one element (nonzero byte = negative)
template <int NQP>
__device__ __forceinline__ uint64_t neg_mask(const uint8_t *s) {
    if constexpr (NQP == 8) return *reinterpret_cast<const uint64_t *>(s);
    else if constexpr (NQP == 4) return *reinterpret_cast<const uint32_t *>(s);
    else {
        uint64_t m = 0;
        for (int q = 0; q < NQP; q++) m |= (uint64_t)s[q] << (8 * q);
        return m;
    }
}

// Adds M^e for the node pair (row ar, column ac) of halo element h to M: the
// sum over the points of s_q u_ar u_ac^T (grad forms; 9 values, or its trace
// for SWP 1) or of N_ar N_ac J w (DOT).  DIAG: a diagonal block (ar == ac),
// whose M^e is symmetric: one vector per point is read and, for SWP 9, only
// the upper triangle (M[0..2], M[4], M[5], M[8]) is formed.
template <int FORM>
constexpr int form_swp() {
    return (FORM == FEASM_FORM_LIN_ELASTIC || FORM == kFormElasticCubic || FORM == FEASM_FORM_DIFFUSION_MAT) ? 9 : 1;
}

// The 16-byte chunk type of the scaled-gradient records (see
// TileShape::uchunks): two doubles or four floats.
template <typename R>
struct Chunk;
template <>
struct Chunk<double> {
    using T = double2;
};
template <>
struct Chunk<float> {
    using T = float4;
};

// s_q r c^T (SWP 9: the 9 products; DIAG: the upper triangle) or its trace
template <int SWP, bool SIGNED, bool DIAG, typename R, typename A>
__device__ __forceinline__ void point_product(bool negative, R c0, R c1, R c2, R r0, R r1, R r2, A *M) {
    if (SIGNED && negative) {
        r0 = -r0;
        r1 = -r1;
        r2 = -r2;
    }
    if (SWP == 9) {
        M[0] = fma(r0, c0, M[0]); M[1] = fma(r0, c1, M[1]); M[2] = fma(r0, c2, M[2]);
        if (!DIAG) M[3] = fma(r1, c0, M[3]);
        M[4] = fma(r1, c1, M[4]); M[5] = fma(r1, c2, M[5]);
        if (!DIAG) { M[6] = fma(r2, c0, M[6]); M[7] = fma(r2, c1, M[7]); }
        M[8] = fma(r2, c2, M[8]);
    } else {
        M[0] = fma(r0, c0, M[0]);
        M[0] = fma(r1, c1, M[0]);
        M[0] = fma(r2, c2, M[0]);
    }
}

template <int NPE, int NQP, int FORM, typename R, bool SIGNED, bool DIAG, typename A>
__device__ __forceinline__ void elem_block(uint32_t it, const R *__restrict__ sU, const uint8_t *__restrict__ sSg,
                                           const double *__restrict__ sN, const R *__restrict__ sW, A *M) {
    constexpr int SWP = form_swp<FORM>();
    const int h = (int)(it >> 8), ac = (int)((it >> 4) & 15u), ar = (int)(it & 15u);
    if constexpr (FORM != FEASM_FORM_DOT) {
        using C = typename Chunk<R>::T;
        constexpr int PX = (sizeof(R) == 8) ? 1 : 2;  // points per xy chunk
        constexpr int PZ = 2 * PX;                    // points per z chunk
        constexpr int US = (NQP * NPE * 3 * (int)sizeof(R) / 16 + 1) * 16 / (int)sizeof(R);  // TileShape::ustride()
        static_assert(NQP % PZ == 0, "points per element must fill the z chunks");
        const C *U = reinterpret_cast<const C *>(sU + h * US);
        const C *UZ = U + (NQP / PX) * NPE;
        uint64_t neg = 0;
        if (SIGNED) neg = neg_mask<NQP>(sSg + h * NQP);
#pragma unroll
        for (int g = 0; g < NQP / PZ; g++) {
            const C cz = UZ[g * NPE + ac];
            const C rz = DIAG ? cz : UZ[g * NPE + ar];
#pragma unroll
            for (int x = 0; x < PZ / PX; x++) {
                const C cxy = U[(g * (PZ / PX) + x) * NPE + ac];
                const C rxy = DIAG ? cxy : U[(g * (PZ / PX) + x) * NPE + ar];
                if constexpr (sizeof(R) == 8) {  // one point: (u0, u1), u2 in cz.x / cz.y
                    const int q = g * 2 + x;
                    const R c2 = x ? cz.y : cz.x, r2 = x ? rz.y : rz.x;
                    point_product<SWP, SIGNED, DIAG>(((neg >> (8 * q)) & 255u) != 0, (R)cxy.x, (R)cxy.y, c2,
                                                     (R)rxy.x, (R)rxy.y, r2, M);
                } else {  // two points: (u0, u1, u0', u1'), u2 of four points in cz
                    const int q = g * 4 + x * 2;
                    const R c2a = x ? cz.z : cz.x, c2b = x ? cz.w : cz.y;
                    const R r2a = x ? rz.z : rz.x, r2b = x ? rz.w : rz.y;
                    point_product<SWP, SIGNED, DIAG>(((neg >> (8 * q)) & 255u) != 0, (R)cxy.x, (R)cxy.y, c2a,
                                                     (R)rxy.x, (R)rxy.y, r2a, M);
                    point_product<SWP, SIGNED, DIAG>(((neg >> (8 * (q + 1))) & 255u) != 0, (R)cxy.z, (R)cxy.w,
                                                     c2b, (R)rxy.z, (R)rxy.w, r2b, M);
                }
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < NQP; q++) {
            const R nn = (R)(sN[q * NPE + ar] * sN[q * NPE + ac]);
            M[0] = fma(nn, sW[h * NQP + q], M[0]);
        }
    }
}

// The element matrices of one block's shared elements, added in ascending
// element order (items j = 0..k-1); ILP elements (1 or 2) in flight at a
// time.  In FP64 the products go straight into the FP64 sums (element after
// element, point after point); in mixed precision each element's matrix is
// formed in FP32 and added in FP64.
template <int NPE, int NQP, int FORM, typename R, bool SIGNED, bool DIAG, int SWP, int ILP>
__device__ __forceinline__ void block_sum(int k, const uint16_t *__restrict__ items, const R *__restrict__ sU,
                                          const uint8_t *__restrict__ sSg, const double *__restrict__ sN,
                                          const R *__restrict__ sW, double *acc) {
    // the next item's index is read while the current one is summed
    if constexpr (sizeof(R) == 8) {
        uint32_t it = items[0];  // k >= 1: every block of the pattern has an element
        for (int j = 0; j < k; j++) {
            const uint32_t nx = items[min(j + 1, k - 1)];
            elem_block<NPE, NQP, FORM, R, SIGNED, DIAG>(it, sU, sSg, sN, sW, acc);
            it = nx;
        }
    } else {
        int j = 0;
        for (; ILP == 2 && j + 1 < k; j += 2) {
            const uint32_t i0 = items[j], i1 = items[j + 1];
            R M0[SWP], M1[SWP];
#pragma unroll
            for (int i = 0; i < SWP; i++) M0[i] = M1[i] = (R)0;
            elem_block<NPE, NQP, FORM, R, SIGNED, DIAG>(i0, sU, sSg, sN, sW, M0);
            elem_block<NPE, NQP, FORM, R, SIGNED, DIAG>(i1, sU, sSg, sN, sW, M1);
#pragma unroll
            for (int i = 0; i < SWP; i++) acc[i] = (acc[i] + (double)M0[i]) + (double)M1[i];
        }
        for (; j < k; j++) {
            R M0[SWP];
#pragma unroll
            for (int i = 0; i < SWP; i++) M0[i] = (R)0;
            elem_block<NPE, NQP, FORM, R, SIGNED, DIAG>(items[j], sU, sSg, sN, sW, M0);
#pragma unroll
            for (int i = 0; i < SWP; i++) acc[i] = acc[i] + (double)M0[i];
        }
    }
}

// Global loads of one tile's LDS image, issued ahead and held in registers:
// NC 16-byte chunks per thread (chunks tid, tid + threads, ...; the rest of a
// larger blob is copied at store time).  COEF: the tile also brings its
// mesh's material coefficients (batched launches); otherwise the launch's
// coefficients are loaded once.
template <bool COEF, int NC>
struct Staged {
    uint4 c[NC];
    double d;  // material coefficient tid (COEF)
    // waits for the loads here: called before the wavefront's global stores,
    // so that the wait does not include them (the vector memory counter
    // counts loads and stores in issue order)
    __device__ __forceinline__ void ready() const {
#pragma unroll
        for (int i = 0; i < NC; i++) asm volatile("" ::"v"(c[i].x), "v"(c[i].y), "v"(c[i].z), "v"(c[i].w));
        if constexpr (COEF) asm volatile("" ::"v"(d));
    }
};
// chunks prefetched per thread (measured: the FP64 and DOT kernels gain from
// a second chunk on the larger T10 blobs and nothing from a third; the
// mixed-precision gradient kernels are faster with one, and so is FP64 scalar
// diffusion, where the compiler spilled a second chunk around every task)
template <typename R, int FORM>
constexpr int staged_chunks() {
    return ((sizeof(R) == 4 && FORM != FEASM_FORM_DOT) || FORM == FEASM_FORM_DIFFUSION_ISO) ? 1 : 2;
}

// A tile header.  With `hdr` a __restrict__ kernel argument (nothing the
// kernel stores aliases it) and t the same in every lane, the compiler reads
// it with scalar loads, so that it sits in scalar registers and no vector
// memory wait is needed for it.
__device__ __forceinline__ TileHdr load_hdr(const TileHdr *__restrict__ hdr, int64_t t) { return hdr[t]; }

template <bool COEF, int NC>
__device__ __forceinline__ void stage_issue(const TileJob &job, const TileHdr &hd, const double *coef,
                                            Staged<COEF, NC> &st) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int nch = hd.nbytes >> 4;  // at least one chunk: every tile owns a column
    const uint4 *src = reinterpret_cast<const uint4 *>(job.blob + hd.blob);
#pragma unroll
    for (int i = 0; i < NC; i++) st.c[i] = src[min(tid + i * nt, nch - 1)];
    if constexpr (COEF) st.d = coef[min(tid, 35)];
}

// The staged image into LDS (blob with the halo coordinates, material
// coefficients); clears the tile's sign flag.
template <bool COEF, int NC>
__device__ __forceinline__ void stage_store(const TileJob &job, const TileHdr &hd, const Staged<COEF, NC> &st,
                                            const TileShape::Layout &L, unsigned char *smem, double *sD,
                                            int *sNeg) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int nch = hd.nbytes >> 4;
    uint4 *dst = reinterpret_cast<uint4 *>(smem + L.cbase);
#pragma unroll
    for (int i = 0; i < NC; i++)
        if (tid + i * nt < nch) dst[tid + i * nt] = st.c[i];
    const uint4 *src = reinterpret_cast<const uint4 *>(job.blob + hd.blob);
    for (int i = tid + NC * nt; i < nch; i += nt) dst[i] = src[i];
    if constexpr (COEF)
        if (tid < 36) sD[tid] = st.d;
    if (tid == 0) *sNeg = 0;
}

// One tile, phases 1 and 2, with its image already in LDS.  `after_phase1`
// runs between the phases; `before_stores` in phase 2, before the first
// global store of every thread.  LDSCOEF: the scalar material coefficients
// are read from LDS (batched launches, where each tile brings its own)
// rather than from the launch's parameters.
template <int NPE, int NQP, int FORM, int NDN, typename R, bool LDSCOEF, typename AfterPhase1,
          typename BeforeStores>
__device__ __forceinline__ void tile_compute(const TileShape &shape, const TileJob &job, const FusedParams &kp,
                                             const TileHdr &hd, unsigned char *smem, const AfterPhase1 &after_phase1,
                                             const BeforeStores &before_stores) {
    constexpr bool GRAD = (FORM != FEASM_FORM_DOT);
    constexpr int SWP = form_swp<FORM>();
    const int tid = threadIdx.x;
    const int kThreads = blockDim.x;
    const int P = hd.P, H = hd.H, NH = hd.NH, NT = hd.NT, K = hd.K, NI = hd.NI;
    constexpr int US = (NQP * NPE * 3 * (int)sizeof(R) / 16 + 1) * 16 / (int)sizeof(R);  // TileShape::ustride()
    // two elements in flight where the registers allow it under 1024 threads
    constexpr int ILP = 2;
    constexpr int DS = NPE * 3;  // TileShape::dstride()
    const TileShape::Layout L = shape.layout(H, P, NH, K, NT, NI);
    double *sN = reinterpret_cast<double *>(smem + L.tab);
    double *sDN = sN + NQP * NPE;
    double *sD = sDN + NQP * DS;  // material coefficients (read as broadcasts)
    int *sNeg = reinterpret_cast<int *>(sD + 36);  // set when J w < 0 somewhere in the tile
    const double *sWq = sD + 37;                   // quadrature weights
    const double *coef = LDSCOEF ? sD : kp.fp.coef;
    const uint64_t *sTask = reinterpret_cast<const uint64_t *>(smem + L.task);
    uint16_t *sItem = reinterpret_cast<uint16_t *>(smem + L.item);
    double *sX = reinterpret_cast<double *>(smem + L.X);
    uint16_t *sLc = reinterpret_cast<uint16_t *>(smem + L.lconn);
    int64_t *sCb = reinterpret_cast<int64_t *>(smem + L.cbase);
    int32_t *sCd = reinterpret_cast<int32_t *>(smem + L.cdeg);
    R *sU = reinterpret_cast<R *>(smem + L.U);
    uint8_t *sSg = reinterpret_cast<uint8_t *>(smem + L.sgn);
    R *sW = reinterpret_cast<R *>(smem + L.sgn);

    // ---- phase 1: Jacobians, inverses and scaled gradients
    for (int task = tid; task < H * NQP; task += kThreads) {
        const int h = task / NQP, q = task % NQP;
        const uint16_t *cn = sLc + h * NPE;
        // coordinates: (x, y) pairs then z values; gradient table: (d0, d1)
        // pairs then d2 values; nodes visited in an order rotated by the
        // point, so that the reads are single 16- and 8-byte loads and the
        // lanes of one element fall into different banks
        const double2 *sXY = reinterpret_cast<const double2 *>(sX);
        const double *sZ = sX + 2 * NH;
        const double2 *dxy = reinterpret_cast<const double2 *>(sDN) + q * NPE;
        const double *dz = sDN + 2 * NQP * NPE + q * NPE;
        double J[9];
        // FP64: the point's gradient table, read once (the stores of the
        // scaled gradients below could alias it, so a second read would be
        // reloaded); measured no faster for FP32, whose kernels keep the reload
        constexpr bool TQ = sizeof(R) == 8;
        double tq[TQ ? 3 * NPE : 1];
#pragma unroll
        for (int i = 0; i < 9; i++) J[i] = 0.0;
#pragma unroll
        for (int i = 0; i < NPE; i++) {
            const int a = (i + q) % NPE;
            const int n = (int)cn[a];
            const double2 xy = sXY[n];
            const double x0 = xy.x, x1 = xy.y, x2 = sZ[n];
            const double2 dd = dxy[a];
            const double d0 = dd.x, d1 = dd.y, d2 = dz[a];
            if constexpr (TQ) {
                tq[3 * i] = d0;
                tq[3 * i + 1] = d1;
                tq[3 * i + 2] = d2;
            }
            J[0] += x0 * d0; J[1] += x0 * d1; J[2] += x0 * d2;
            J[3] += x1 * d0; J[4] += x1 * d1; J[5] += x1 * d2;
            J[6] += x2 * d0; J[7] += x2 * d1; J[8] += x2 * d2;
        }
#define RJ(i, j) J[3 * ((i)-1) + ((j)-1)]
        const double det = RJ(1, 1) * (RJ(2, 2) * RJ(3, 3) - RJ(3, 2) * RJ(2, 3)) -
                           RJ(1, 2) * (RJ(2, 1) * RJ(3, 3) - RJ(2, 3) * RJ(3, 1)) +
                           RJ(1, 3) * (RJ(2, 1) * RJ(3, 2) - RJ(2, 2) * RJ(3, 1));
        double wj = (FORM == FEASM_FORM_DIFFUSION_ISO) ? (coef[0] * det) * sWq[q] : det * sWq[q];
        if (GRAD) {
            // the inverse scaled by sqrt(|J w|) once: g_a scaled is inv(J)' dN_a times it
            const double invdet = (1.0 / det) * sqrt(fabs(wj));
            const double iv[9] = {(RJ(2, 2) * RJ(3, 3) - RJ(3, 2) * RJ(2, 3)) * invdet,
                                  -(RJ(1, 2) * RJ(3, 3) - RJ(1, 3) * RJ(3, 2)) * invdet,
                                  (RJ(1, 2) * RJ(2, 3) - RJ(1, 3) * RJ(2, 2)) * invdet,
                                  -(RJ(2, 1) * RJ(3, 3) - RJ(2, 3) * RJ(3, 1)) * invdet,
                                  (RJ(1, 1) * RJ(3, 3) - RJ(1, 3) * RJ(3, 1)) * invdet,
                                  -(RJ(1, 1) * RJ(2, 3) - RJ(2, 1) * RJ(1, 3)) * invdet,
                                  (RJ(2, 1) * RJ(3, 2) - RJ(3, 1) * RJ(2, 2)) * invdet,
                                  -(RJ(1, 1) * RJ(3, 2) - RJ(3, 1) * RJ(1, 2)) * invdet,
                                  (RJ(1, 1) * RJ(2, 2) - RJ(2, 1) * RJ(1, 2)) * invdet};
            sSg[h * NQP + q] = (wj < 0.0) ? 1 : 0;
            if (wj < 0.0) atomicOr(sNeg, 1);
            R ivr[9];
#pragma unroll
            for (int i = 0; i < 9; i++) ivr[i] = (R)iv[i];
            // the element's record of 16-byte chunks (TileShape::uchunks):
            // (u0, u1) of point q in xy chunk q / PX, u2 in z chunk q / PZ
            constexpr int PX = (sizeof(R) == 8) ? 1 : 2, PZ = 2 * PX, CR = 16 / (int)sizeof(R);
            R *urec = sU + (size_t)h * US;
            // node order rotated by the point, so that the lanes of one
            // element store into different LDS banks
#pragma unroll
            for (int i = 0; i < NPE; i++) {
                const int a = (i + q) % NPE;
                R d0, d1, d2;
                if constexpr (TQ) {
                    d0 = tq[3 * i];
                    d1 = tq[3 * i + 1];
                    d2 = tq[3 * i + 2];
                } else {
                    const double2 dd = dxy[a];
                    d0 = (R)dd.x;
                    d1 = (R)dd.y;
                    d2 = (R)dz[a];
                }
                const R u0 = d0 * ivr[0] + d1 * ivr[3] + d2 * ivr[6];
                const R u1 = d0 * ivr[1] + d1 * ivr[4] + d2 * ivr[7];
                const R u2 = d0 * ivr[2] + d1 * ivr[5] + d2 * ivr[8];
                const int cxy = (q / PX) * NPE + a, cz = (NQP / PX) * NPE + (q / PZ) * NPE + a;
                if constexpr (sizeof(R) == 8)
                    reinterpret_cast<double2 *>(urec)[cxy] = make_double2(u0, u1);
                else
                    reinterpret_cast<float2 *>(urec)[cxy * 2 + (q % PX)] = make_float2(u0, u1);
                urec[cz * CR + (q % PZ)] = u2;
            }
        } else {
            sW[h * NQP + q] = (R)wj;
        }
#undef RJ
    }
    // one flag for the tile: J w < 0 somewhere (inverted elements) selects
    // the path that applies the signs
    __syncthreads();
    const bool tile_neg = *sNeg != 0;
    after_phase1();


    // ---- phase 2: one thread per node block: element matrices of the shared
    // elements, ordered sum, material, one write
    double *__restrict__ nzval = job.nzval;
    // one block task: its values and where they go
    struct BlockOut {
        double out[NDN * NDN];
        int64_t base, base2;
        int deg, deg2;
        bool valid, pair;
    };
    const auto compute = [&](int t, BlockOut &b) {
        b.valid = false;
        const uint64_t rec = sTask[t];
        const uint32_t th = (uint32_t)rec;
        if (th == kNullTask) return;
        const int c = (int)(th & 255u), r = (int)((th >> 8) & 255u), k = (int)((th >> 16) & 255u);
        const bool sw = (th >> 24) & 1u, diag = (th >> 25) & 1u, sym = (th >> 26) & 1u;
        double acc[SWP];
#pragma unroll
        for (int i = 0; i < SWP; i++) acc[i] = 0.0;
        const uint16_t *items = sItem + (uint32_t)((rec >> 32) & 0xffffu);
        if (GRAD && sym) {  // symmetric diagonal blocks sit in wavefronts of their own
            if (tile_neg)
                block_sum<NPE, NQP, FORM, R, true, true, SWP, ILP>(k, items, sU, sSg, sN, sW, acc);
            else
                block_sum<NPE, NQP, FORM, R, false, true, SWP, ILP>(k, items, sU, sSg, sN, sW, acc);
            if constexpr (SWP == 9) {  // the lower triangle of the symmetric sum
                acc[3] = acc[1];
                acc[6] = acc[2];
                acc[7] = acc[5];
            }
        } else {
            if (tile_neg)
                block_sum<NPE, NQP, FORM, R, true, false, SWP, ILP>(k, items, sU, sSg, sN, sW, acc);
            else
                block_sum<NPE, NQP, FORM, R, false, false, SWP, ILP>(k, items, sU, sSg, sN, sW, acc);
        }
        double *out = b.out;
        if constexpr (SWP == 9) {
            // contract in the orientation of the lower node number
            double Mc[9];
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++) Mc[3 * a + b] = sw ? acc[3 * b + a] : acc[3 * a + b];
            if constexpr (FORM == FEASM_FORM_DIFFUSION_MAT) {
                // the address depends on the task, which keeps the compiler
                // from holding every coefficient in registers across tasks
                const double *D = sD + (th >> 31);
                double s = 0.0;
#pragma unroll
                for (int a = 0; a < 3; a++)
#pragma unroll
                    for (int b = 0; b < 3; b++) s += D[3 * a + b] * Mc[3 * a + b];
                out[0] = s;
            } else {
                double Kc[9];
                if constexpr (FORM == kFormElasticCubic) {
                    const double d00 = coef[0], d01 = coef[1], d33 = coef[21];
                    const double tr = (Mc[0] + Mc[4]) + Mc[8];
#pragma unroll
                    for (int i = 0; i < 3; i++)
#pragma unroll
                        for (int jj = 0; jj < 3; jj++)
                            Kc[3 * i + jj] = (i == jj) ? d00 * Mc[4 * i] + d33 * (tr - Mc[4 * i])
                                                       : d01 * Mc[3 * i + jj] + d33 * Mc[3 * jj + i];
                } else {
                    const double *D = sD + (th >> 31);  // see above
#pragma unroll
                    for (int i = 0; i < 3; i++)
#pragma unroll
                        for (int jj = 0; jj < 3; jj++) {
                            double s = 0.0;
#pragma unroll
                            for (int a = 0; a < 3; a++)
#pragma unroll
                                for (int b = 0; b < 3; b++) s += D[6 * voigt(i, a) + voigt(jj, b)] * Mc[3 * a + b];
                            Kc[3 * i + jj] = s;
                        }
                }
                if (diag) {  // diagonal block: upper triangle mirrored, as complete_lt! does
                    Kc[3] = Kc[1];
                    Kc[6] = Kc[2];
                    Kc[7] = Kc[5];
                }
#pragma unroll
                for (int a = 0; a < 3; a++)
#pragma unroll
                    for (int b = 0; b < 3; b++) out[(NDN * a + b) % (NDN * NDN)] = sw ? Kc[3 * b + a] : Kc[3 * a + b];
            }
        } else if constexpr (FORM == FEASM_FORM_DOT) {
            const double *D = sD + (th >> 31);  // see above
#pragma unroll
            for (int a = 0; a < NDN; a++)
#pragma unroll
                for (int b = 0; b < NDN; b++) out[NDN * a + b] = acc[0] * D[NDN * a + b];
        } else {
            out[0] = acc[0];
        }
        b.deg = sCd[c];
        b.base = (int64_t)NDN * NDN * sCb[c] + (int64_t)r * NDN;
        const uint32_t pr = (uint32_t)(rec >> 48);
        b.pair = pr != kNoPair;
        if (b.pair) {  // the transposed block (row c, column m), in its own column
            const int c2 = (int)(pr & 255u), r2 = (int)(pr >> 8);
            b.deg2 = sCd[c2];
            b.base2 = (int64_t)NDN * NDN * sCb[c2] + (int64_t)r2 * NDN;
        }
        b.valid = true;
    };
    const auto store = [&](const BlockOut &b) {
        if (!b.valid) return;
#pragma unroll
        for (int l = 0; l < NDN; l++)
#pragma unroll
            for (int a = 0; a < NDN; a++) nzval[b.base + (int64_t)l * NDN * b.deg + a] = b.out[NDN * a + l];
        if (b.pair) {
            // K_cm = K_mc^T; the DOT forms scale a scalar by the same matrix
            // in both blocks
#pragma unroll
            for (int l = 0; l < NDN; l++)
#pragma unroll
                for (int a = 0; a < NDN; a++)
                    nzval[b.base2 + (int64_t)l * NDN * b.deg2 + a] =
                        (FORM == FEASM_FORM_DOT) ? b.out[NDN * a + l] : b.out[NDN * l + a];
        }
    };
    // The first task of every thread is computed before `before_stores`
    // runs and stored after it: the wait there (for loads issued before
    // phase 1) then comes as late as it can without including any store.
    {
        BlockOut b;
        b.valid = false;
        if (tid < NT) compute(tid, b);
        before_stores();
        store(b);
    }
    for (int t = tid + kThreads; t < NT; t += kThreads) {
        BlockOut b;
        compute(t, b);
        store(b);
    }
    __syncthreads();  // LDS is reused by the workgroup's next tile
}

// The tables of the element family, once per workgroup.
template <int NPE, int NQP>
__device__ __forceinline__ void load_tables(const FusedParams &kp, unsigned char *smem) {
    double *sN = reinterpret_cast<double *>(smem);  // the layout's tab region starts at 0
    double *sDN = sN + NQP * NPE;  // (d0, d1) pairs [q][a], then d2 [q][a]
    double *sWq = sDN + NQP * NPE * 3 + 37;  // after the material coefficients and the flag
    for (int i = threadIdx.x; i < NQP; i += blockDim.x) sWq[i] = kp.w[i];
    for (int i = threadIdx.x; i < NQP * NPE; i += blockDim.x) sN[i] = kp.N[i];
    for (int i = threadIdx.x; i < NQP * NPE; i += blockDim.x) {
        sDN[2 * i] = kp.dN[3 * i];
        sDN[2 * i + 1] = kp.dN[3 * i + 1];
        sDN[2 * NQP * NPE + i] = kp.dN[3 * i + 2];
    }
}

// XCD-aware persistent walk: consecutive workgroups land on different XCDs,
// so each XCD takes a contiguous range of the (Morton-ordered) tile list and
// its workgroups stride through it.
struct TileRange {
    int64_t begin, end, step;
};
__device__ __forceinline__ TileRange xcd_range(int64_t ntiles) {
    const int64_t per = (ntiles + 7) / 8;
    const int64_t xcd = blockIdx.x % 8, slot = blockIdx.x / 8, nslots = gridDim.x / 8;
    const int64_t b = xcd * per;
    const int64_t e = (b + per < ntiles) ? b + per : ntiles;
    return TileRange{b + slot, e, nslots};
}

template <int NPE, int NQP, int FORM, int NDN, typename R>
__global__ __launch_bounds__(kMaxThreads) void k_fused(TileShape shape, TileJob job,
                                                    const TileHdr *__restrict__ hdr, int64_t ntiles,
                                                    FusedParams kp) {
    extern __shared__ __align__(16) unsigned char smem[];
    const TileRange rg = xcd_range(ntiles);
    if (rg.begin >= rg.end) return;
    constexpr int DS = NPE * 3;  // TileShape::dstride()
    double *sD = reinterpret_cast<double *>(smem) + NQP * NPE + NQP * DS;
    int *sNeg = reinterpret_cast<int *>(sD + 36);
    load_tables<NPE, NQP>(kp, smem);
    for (int i = threadIdx.x; i < 36; i += blockDim.x) sD[i] = kp.fp.coef[i];  // one material per launch
    // Tile t runs while the blob of tile t1 (the next one) is loaded and
    // the header of tile t2 is read.  For the gradient forms the blob load is
    // issued before phase 1 and waited for before phase 2's global stores:
    // the vector memory counter counts loads and stores together, in issue
    // order, so a load waited for after the stores would wait for every
    // store too.  The DOT forms have too short a phase 1 to hide the load:
    // they issue it after phase 1 and wait for it (and the stores) at the
    // next tile (measured faster for them).
    constexpr bool EARLY = (FORM != FEASM_FORM_DOT);
    int64_t t1 = rg.begin + rg.step;
    TileHdr hd = load_hdr(hdr, rg.begin);
    TileHdr hn = load_hdr(hdr, t1 < rg.end ? t1 : rg.begin);
    // (the loads are unconditional, from valid tiles, so that no register
    // copy of a pending load forces an early wait)
    Staged<false, staged_chunks<R, FORM>()> st;
    stage_issue(job, hd, nullptr, st);
    st.ready();
    for (;;) {
        const bool more = t1 < rg.end;
        const int64_t t2 = t1 + rg.step;
        const int64_t i2 = t2 < rg.end ? t2 : rg.begin;
        // the gradient forms read the header of tile t2 while the blob goes
        // to LDS: the barrier's wait covers it, and no scalar load is left
        // pending in the phases (that would turn every wait for an LDS read
        // into a wait for all of them); the store-bound DOT forms measured
        // faster with the read after the barrier
        TileHdr h2;
        if constexpr (EARLY) h2 = load_hdr(hdr, i2);
        stage_store(job, hd, st, shape.layout(hd.H, hd.P, hd.NH, hd.K, hd.NT, hd.NI), smem, sD, sNeg);
        __syncthreads();
        if constexpr (!EARLY) h2 = load_hdr(hdr, i2);
        const TileHdr &hb = more ? hn : hd;
        if constexpr (EARLY) stage_issue(job, hb, nullptr, st);
        tile_compute<NPE, NQP, FORM, NDN, R, false>(
            shape, job, kp, hd, smem,
            [&]() {
                if constexpr (!EARLY) stage_issue(job, hb, nullptr, st);
            },
            [&]() {
                if constexpr (EARLY) st.ready();
            });
        if (!more) break;
        t1 = t2;
        hd = hn;
        hn = h2;
    }
}

// One entry per mesh of a batched launch.
struct BatchEntry {
    TileJob job;
    FusedParams kp;
    TileShape shape;
    int64_t tile0;
};

// Persistent batched launch: as k_fused over the global tile list (all
// meshes' tiles, mesh after mesh); every tile carries its mesh's
// coefficients and arrays.  One launch covers the meshes of one element
// type, rule, arithmetic and form.
template <int NPE, int NQP, int FORM, int NDN, typename R>
__global__ __launch_bounds__(kMaxThreads) void k_fused_batch(const BatchEntry *__restrict__ entries,
                                                          const int32_t *__restrict__ tile_entry,
                                                          const TileHdr *__restrict__ hdr, int64_t ntiles,
                                                          TileShape shape) {
    extern __shared__ __align__(16) unsigned char smem[];
    const TileRange rg = xcd_range(ntiles);
    if (rg.begin >= rg.end) return;
    constexpr int DS = NPE * 3;  // TileShape::dstride()
    double *sD = reinterpret_cast<double *>(smem) + NQP * NPE + NQP * DS;
    int *sNeg = reinterpret_cast<int *>(sD + 36);
    load_tables<NPE, NQP>(entries[tile_entry[rg.begin]].kp, smem);
    // the tile walk and its loads as in k_fused; every tile also brings its
    // mesh's coefficients
    constexpr bool EARLY = (FORM != FEASM_FORM_DOT);
    int64_t t1 = rg.begin + rg.step;
    TileHdr hd = load_hdr(hdr, rg.begin);
    const BatchEntry *E = entries + tile_entry[rg.begin];
    const int64_t i1 = t1 < rg.end ? t1 : rg.begin;
    TileHdr hn = load_hdr(hdr, i1);
    const BatchEntry *En = entries + tile_entry[i1];
    Staged<true, staged_chunks<R, FORM>()> st;
    stage_issue(E->job, hd, E->kp.fp.coef, st);
    st.ready();
    for (;;) {
        const bool more = t1 < rg.end;
        const int64_t t2 = t1 + rg.step;
        const int64_t i2 = t2 < rg.end ? t2 : rg.begin;
        TileHdr h2;
        const BatchEntry *E2 = nullptr;
        if constexpr (EARLY) {
            h2 = load_hdr(hdr, i2);
            E2 = entries + tile_entry[i2];
        }
        stage_store(E->job, hd, st, shape.layout(hd.H, hd.P, hd.NH, hd.K, hd.NT, hd.NI), smem, sD, sNeg);
        __syncthreads();
        if constexpr (!EARLY) {
            h2 = load_hdr(hdr, i2);
            E2 = entries + tile_entry[i2];
        }
        const TileHdr &hb = more ? hn : hd;
        const BatchEntry *Eb = more ? En : E;
        if constexpr (EARLY) stage_issue(Eb->job, hb, Eb->kp.fp.coef, st);
        tile_compute<NPE, NQP, FORM, NDN, R, true>(
            shape, E->job, E->kp, hd, smem,
            [&]() {
                if constexpr (!EARLY) stage_issue(Eb->job, hb, Eb->kp.fp.coef, st);
            },
            [&]() {
                if constexpr (EARLY) st.ready();
            });
        if (!more) break;
        t1 = t2;
        hd = hn;
        E = En;
        hn = h2;
        En = E2;
    }
}

TileJob make_job(const feasm_mesh *m, const TileSet &t, double *nzval) {
    TileJob j;
    j.hdr = static_cast<const TileHdr *>(t.hdr->p);
    j.blob = static_cast<const unsigned char *>(t.blob->p);
    j.hnodes = dp<int32_t>(t.hnodes);
    j.xyz = dp<double>(m->xyz);
    j.nzval = nzval;
    return j;
}

// workgroups of a persistent launch: as many as stay resident (LDS-bound),
// a multiple of the 8 XCDs, and no more than the tiles need
unsigned persistent_grid(feasm_ctx *ctx, size_t budget, int64_t ntiles) {
    int cus = 0;
    FEASM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const size_t lds_cu = 160 * 1024;
    const int per_cu = (int)std::max<size_t>(1, lds_cu / std::max<size_t>(budget, 1));
    int64_t grid = (int64_t)cus * per_cu;
    grid = std::max<int64_t>(8, (grid / 8) * 8);
    const int64_t need = 8 * ((ntiles + 7) / 8);
    return (unsigned)std::min(grid, std::max<int64_t>(need, 8));
}

template <int NPE, int NQP, int FORM, int NDN, typename R>
void go(feasm_ctx *ctx, const feasm_mesh *m, const Pattern &p, const TileSet &t, const FusedParams &kp,
        double *nzval) {
    auto kern = k_fused<NPE, NQP, FORM, NDN, R>;
    const int budget = (int)t.shape.budget;
    FEASM_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, budget));
    const unsigned grid = persistent_grid(ctx, t.shape.budget, t.ntiles);
    TileJob job = make_job(m, t, nzval);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(t.shape.threads), budget, ctx->stream, t.shape, job, job.hdr, t.ntiles,
                       kp);
    FEASM_HIP(hipGetLastError());
}

template <int NPE, int NQP, typename R>
void by_form(feasm_ctx *ctx, const feasm_mesh *m, const Pattern &p, const TileSet &t, const FusedParams &kp,
             double *nzval) {
    switch (kp.fp.form) {
        case FEASM_FORM_LIN_ELASTIC:
            if (kp.cubic) return go<NPE, NQP, kFormElasticCubic, 3, R>(ctx, m, p, t, kp, nzval);
            return go<NPE, NQP, FEASM_FORM_LIN_ELASTIC, 3, R>(ctx, m, p, t, kp, nzval);
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

template <int NPE, int NQP, int FORM, int NDN, typename R>
void go_batch(feasm_ctx *ctx, const BatchLaunch &b) {
    auto kern = k_fused_batch<NPE, NQP, FORM, NDN, R>;
    FEASM_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b.budget));
    hipLaunchKernelGGL(kern, dim3(b.grid), dim3(b.threads), b.budget, ctx->stream,
                       static_cast<const BatchEntry *>(b.entries->p), dp<int32_t>(b.tile_entry),
                       static_cast<const TileHdr *>(b.hdr->p), b.ntiles, b.shape);
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
        E.job = make_job(it.mesh, *it.tiles, it.nzval);
        E.kp.fp = it.fp;
        E.kp.cubic = (it.fp.form == FEASM_FORM_LIN_ELASTIC && cubic_symmetry(it.fp.coef)) ? 1 : 0;
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
        if (i == 0) b.shape = it.tiles->shape;
        const TileShape &s0 = b.shape, &si = it.tiles->shape;
        FEASM_REQUIRE(si.npe == s0.npe && si.nqp == s0.nqp && si.rbytes == s0.rbytes && si.grad == s0.grad,
                      FEASM_ESTATE, "meshes of one batched launch differ in their tile layout");
        host.push_back(E);
    }
    b.entries = dev_alloc(std::max<size_t>(host.size() * sizeof(BatchEntry), 16));
    b.tile_entry = dev_alloc(std::max<size_t>(tile_entry.size() * 4, 4));
    b.hdr = dev_alloc(std::max<size_t>((size_t)b.ntiles * sizeof(TileHdr), 16));
    FEASM_HIP(hipMemcpy(b.entries->p, host.data(), host.size() * sizeof(BatchEntry), hipMemcpyHostToDevice));
    FEASM_HIP(hipMemcpy(b.tile_entry->p, tile_entry.data(), tile_entry.size() * 4, hipMemcpyHostToDevice));
    for (size_t i = 0; i < items.size(); i++)  // the meshes' headers, one after the other
        FEASM_HIP(hipMemcpy(static_cast<TileHdr *>(b.hdr->p) + host[i].tile0, items[i].tiles->hdr->p,
                            (size_t)items[i].tiles->ntiles * sizeof(TileHdr), hipMemcpyDeviceToDevice));
    b.grid = persistent_grid(ctx, b.budget, b.ntiles);
}

namespace {

template <int NPE, int NQP, typename R>
void batch_by_form(feasm_ctx *ctx, const BatchLaunch &b) {
    switch (b.kform) {
        case FEASM_FORM_LIN_ELASTIC: return go_batch<NPE, NQP, FEASM_FORM_LIN_ELASTIC, 3, R>(ctx, b);
        case kFormElasticCubic: return go_batch<NPE, NQP, kFormElasticCubic, 3, R>(ctx, b);
        case FEASM_FORM_DIFFUSION_ISO: return go_batch<NPE, NQP, FEASM_FORM_DIFFUSION_ISO, 1, R>(ctx, b);
        case FEASM_FORM_DIFFUSION_MAT: return go_batch<NPE, NQP, FEASM_FORM_DIFFUSION_MAT, 1, R>(ctx, b);
        case FEASM_FORM_DOT:
            if (b.ndn == 1) return go_batch<NPE, NQP, FEASM_FORM_DOT, 1, R>(ctx, b);
            if (b.ndn == 2) return go_batch<NPE, NQP, FEASM_FORM_DOT, 2, R>(ctx, b);
            if (b.ndn == 3) return go_batch<NPE, NQP, FEASM_FORM_DOT, 3, R>(ctx, b);
            break;
    }
    throw Error(FEASM_EUNSUPPORTED, "unsupported form in the batched launch");
}

}  // namespace

int fused_kernel_form(const FormParams &fp) {
    if (fp.form == FEASM_FORM_LIN_ELASTIC && cubic_symmetry(fp.coef)) return kFormElasticCubic;
    return fp.form;
}

void launch_batch(feasm_ctx *ctx, const BatchLaunch &b) {
    if (b.ntiles == 0) return;
    if (b.npe == 8 && b.nqp == 8) {
        if (b.rbytes == 4) return batch_by_form<8, 8, float>(ctx, b);
        return batch_by_form<8, 8, double>(ctx, b);
    }
    if (b.npe == 10 && b.nqp == 4) {
        if (b.rbytes == 4) return batch_by_form<10, 4, float>(ctx, b);
        return batch_by_form<10, 4, double>(ctx, b);
    }
    throw Error(FEASM_EUNSUPPORTED, "batched launch covers H8 2x2x2 Gauss and T10 4-point rules");
}

namespace {

// one workgroup per tile: the tile's halo coordinates into its blob
__global__ __launch_bounds__(256) void k_tile_coords(TileShape shape, const TileHdr *__restrict__ hdr,
                                                     int64_t ntiles, const int32_t *__restrict__ hnodes,
                                                     const double *__restrict__ xyz, unsigned char *__restrict__ blob) {
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const TileHdr hd = hdr[t];
        const TileShape::Layout L = shape.layout(hd.H, hd.P, hd.NH, hd.K, hd.NT, hd.NI);
        double *X = reinterpret_cast<double *>(blob + hd.blob + (L.X - L.cbase));
        for (int i = threadIdx.x; i < hd.NH; i += blockDim.x) {  // (x, y) pairs, then the z values
            const int64_t n = hnodes[hd.x0 + i];
            X[2 * i] = xyz[3 * n];
            X[2 * i + 1] = xyz[3 * n + 1];
            X[2 * hd.NH + i] = xyz[3 * n + 2];
        }
    }
}

}  // namespace

void refresh_tile_coords(feasm_mesh *m, TileSet &t) {
    if (t.ntiles == 0) return;
    feasm_ctx *ctx = m->ctx;
    hipLaunchKernelGGL(k_tile_coords, dim3((unsigned)std::min<int64_t>(t.ntiles, 1 << 16)), dim3(256), 0,
                       ctx->stream, t.shape, static_cast<const TileHdr *>(t.hdr->p), t.ntiles, dp<int32_t>(t.hnodes),
                       dp<double>(m->xyz), static_cast<unsigned char *>(t.blob->p));
    FEASM_HIP(hipGetLastError());
}

bool fused_supported(int npe, int nqp) { return (npe == 8 && nqp == 8) || (npe == 10 && nqp == 4); }

TileShape tile_shape_for(feasm_ctx *ctx, int npe, int nqp, int form, int precision) {
    TileShape s;
    s.npe = npe;
    s.nqp = nqp;
    s.rbytes = (precision == FEASM_MIXED) ? 4 : 8;
    s.grad = (form != FEASM_FORM_DOT);
    int lds = 0;
    FEASM_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
    s.budget = (size_t)std::min(lds, 160 * 1024);
    // measured best on MI355X: DOT tiles (tiny records, not bound by LDS)
    // with 64 nodes list more block tasks than the workgroup has threads
    s.max_nodes = (form == FEASM_FORM_DOT) ? 32 : 64;
    s.max_halo = 255;
    s.threads = 1024;
    return s;
}

void launch_fused(feasm_ctx *ctx, const feasm_mesh *m, const Pattern &p, const TileSet &t, const DevQTable &qt,
                  const FormParams &fp, int precision, double *nzval) {
    if (t.ntiles == 0) return;
    FusedParams kp;
    kp.fp = fp;
    kp.cubic = (fp.form == FEASM_FORM_LIN_ELASTIC && cubic_symmetry(fp.coef)) ? 1 : 0;
    kp.N = qt.N();
    kp.dN = qt.dN();
    kp.w = qt.w();
    if (precision == FEASM_MIXED)
        dispatch<float>(ctx, m, p, t, kp, nzval);
    else
        dispatch<double>(ctx, m, p, t, kp, nzval);
}

}  // namespace feasm
