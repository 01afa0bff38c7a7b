// Measurement helper (no reference counterpart): throughput of the
// per-element node-pair Gram of an 8-node, 8-point element,
//     M_ab = sum_q u_a(q) u_b(q)^T   (3 x 3 per node pair a <= b),
// the product at the core of the element matrices (the upper triangle that
// add_btdb_ut_only! forms, src/MatrixUtilityModule.jl:189-216, before the
// material is applied), computed from element records held in LDS either on
// the FP64 vector ALUs (one lane per (element, node pair), 72 FMA each) or on
// the FP64 matrix cores: the element's 24 x 8 gradient table G (rows 4a + k,
// padded to 32; columns the points) gives G G^T with six
// v_mfma_f64_16x16x4f64 (the three upper 16 x 16 tiles, two 4-point k
// blocks each).  Both read the same records; a batch of 16 elements sits in
// LDS and is reused `reps` times so that the figure is the arithmetic, not
// the loads.  Element values are a hash of (element, point, node, component).
#include "common.hpp"

namespace feasm {

namespace {

typedef double double4_t __attribute__((ext_vector_type(4)));

constexpr int kBatch = 16;   // elements per LDS batch
constexpr int kRec = 8 * 8 * 4;  // doubles per element record: [q][a][k], k padded to 4

__device__ __forceinline__ double uval(int64_t e, int q, int a, int k) {
    uint64_t z = (uint64_t)e * 0x9e3779b97f4a7c15ull + (uint64_t)((q * 8 + a) * 4 + k) * 0xbf58476d1ce4e5b9ull;
    z ^= z >> 31;
    z *= 0x94d049bb133111ebull;
    z ^= z >> 29;
    return (double)(z >> 44) * (1.0 / 1048576.0) - 0.5;
}

// fills LDS with the records of elements [e0, e0 + kBatch)
__device__ __forceinline__ void fill_batch(double *rec, int64_t e0) {
    for (int i = threadIdx.x; i < kBatch * kRec; i += blockDim.x) {
        const int el = i / kRec, r = i % kRec, q = r / 32, a = (r / 4) % 8, k = r % 4;
        rec[i] = (k < 3) ? uval(e0 + el, q, a, k) : 0.0;
    }
}

// pair index p (0..35) -> (a, b), a <= b
__device__ __forceinline__ void pair_ab(int p, int &a, int &b) {
    b = 0;
    while ((b + 1) * (b + 2) / 2 <= p) b++;
    a = p - b * (b + 1) / 2;
}

__global__ __launch_bounds__(256) void k_gram_valu(int64_t nbatch, int reps, double *__restrict__ part,
                                                   double *__restrict__ grams, int64_t nkeep) {
    __shared__ __align__(16) double rec[kBatch * kRec];
    double acc = 0.0;
    for (int64_t bt = blockIdx.x; bt < nbatch; bt += gridDim.x) {
        __syncthreads();
        fill_batch(rec, bt * kBatch);
        __syncthreads();
        for (int r = 0; r < reps; r++)
            for (int t = threadIdx.x; t < kBatch * 36; t += blockDim.x) {
                const int el = t / 36;
                int a, b;
                pair_ab(t % 36, a, b);
                const double4_t *R = reinterpret_cast<const double4_t *>(rec + el * kRec);
                double M[9];
#pragma unroll
                for (int i = 0; i < 9; i++) M[i] = 0.0;
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const double4_t ua = R[q * 8 + a], ub = R[q * 8 + b];
#pragma unroll
                    for (int k = 0; k < 3; k++)
#pragma unroll
                        for (int l = 0; l < 3; l++) M[3 * k + l] = fma(ua[k], ub[l], M[3 * k + l]);
                }
#pragma unroll
                for (int i = 0; i < 9; i++) acc = fma(M[i], (double)(i + 1), acc);
                const int64_t e = bt * kBatch + el;
                if (r == 0 && e < nkeep)
                    for (int i = 0; i < 9; i++) grams[(e * 36 + t % 36) * 9 + i] = M[i];
            }
    }
    part[blockIdx.x * (int64_t)blockDim.x + threadIdx.x] = acc;
}

// v_mfma_f64_16x16x4f64 lane layout (checked on gfx950 against the vector
// variant): lane l holds A[l % 16][l / 16] and B[l / 16][l % 16]; register i
// of the result holds D[4 i + l / 16][l % 16]
__global__ __launch_bounds__(256) void k_gram_mfma(int64_t nbatch, int reps, double *__restrict__ part,
                                                   double *__restrict__ grams, int64_t nkeep) {
    __shared__ __align__(16) double rec[kBatch * kRec];
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    const int row = l % 16, kq = l / 16;  // A operand: row of the tile, point within the k block
    double acc = 0.0;
    for (int64_t bt = blockIdx.x; bt < nbatch; bt += gridDim.x) {
        __syncthreads();
        fill_batch(rec, bt * kBatch);
        __syncthreads();
        for (int r = 0; r < reps; r++)
            for (int el = w; el < kBatch; el += blockDim.x / 64) {
                const double *R = rec + el * kRec;
                double A[2][2];  // [row tile][k block]
#pragma unroll
                for (int t = 0; t < 2; t++)
#pragma unroll
                    for (int kb = 0; kb < 2; kb++) {
                        const int Rr = 16 * t + row, a = Rr / 4, k = Rr % 4, q = 4 * kb + kq;
                        A[t][kb] = R[(q * 8 + a) * 4 + k];  // the pad slot k = 3 holds 0
                    }
                double4_t C00 = {0, 0, 0, 0}, C01 = {0, 0, 0, 0}, C11 = {0, 0, 0, 0};
#pragma unroll
                for (int kb = 0; kb < 2; kb++) {
                    C00 = __builtin_amdgcn_mfma_f64_16x16x4f64(A[0][kb], A[0][kb], C00, 0, 0, 0);
                    C01 = __builtin_amdgcn_mfma_f64_16x16x4f64(A[0][kb], A[1][kb], C01, 0, 0, 0);
                    C11 = __builtin_amdgcn_mfma_f64_16x16x4f64(A[1][kb], A[1][kb], C11, 0, 0, 0);
                }
                const int64_t e = bt * kBatch + el;
                const bool keep = (r == 0 && e < nkeep);
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int drow = 4 * i + l / 16, dcol = l % 16;
                    const double v[3] = {C00[i], C01[i], C11[i]};
#pragma unroll
                    for (int tt = 0; tt < 3; tt++) {
                        const int t1 = tt == 2 ? 1 : 0, t2 = tt == 0 ? 0 : 1;
                        const int Ra = 16 * t1 + drow, Cb = 16 * t2 + dcol;
                        const int a = Ra / 4, k = Ra % 4, b = Cb / 4, ll = Cb % 4;
                        if (k < 3 && ll < 3 && a <= b) {
                            acc = fma(v[tt], (double)(3 * k + ll + 1), acc);
                            if (keep) grams[(e * 36 + b * (b + 1) / 2 + a) * 9 + 3 * k + ll] = v[tt];
                        }
                    }
                }
            }
    }
    part[blockIdx.x * (int64_t)blockDim.x + threadIdx.x] = acc;
}

}  // namespace

}  // namespace feasm

using namespace feasm;

extern "C" int feasm_probe_gram(feasm_ctx *ctx, int variant, int64_t nelem, int reps, double *ms,
                                double *checksum, double *grams_host, int64_t nkeep) {
    try {
        FEASM_REQUIRE(ctx && ms && checksum, FEASM_EINVAL, "null argument");
        FEASM_REQUIRE(variant == 0 || variant == 1, FEASM_EINVAL, "variant: 0 vector FP64, 1 matrix FP64");
        FEASM_REQUIRE(nelem >= kBatch && nelem % kBatch == 0 && reps >= 1, FEASM_EINVAL,
                      "nelem a positive multiple of 16, reps >= 1");
        FEASM_REQUIRE(nkeep >= 0 && nkeep <= nelem && (nkeep == 0 || grams_host), FEASM_EINVAL, "bad nkeep");
        const int64_t nbatch = nelem / kBatch;
        int cus = 0;
        FEASM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
        const unsigned grid = (unsigned)std::min<int64_t>(nbatch, (int64_t)cus * 8);
        DevPtr part = dev_alloc((size_t)grid * 256 * sizeof(double));
        DevPtr grams = dev_alloc(std::max<size_t>((size_t)nkeep * 36 * 9 * sizeof(double), 8));
        hipEvent_t e0, e1;
        FEASM_HIP(hipEventCreate(&e0));
        FEASM_HIP(hipEventCreate(&e1));
        FEASM_HIP(hipEventRecord(e0, ctx->stream));
        if (variant == 0)
            hipLaunchKernelGGL(k_gram_valu, dim3(grid), dim3(256), 0, ctx->stream, nbatch, reps, dp<double>(part),
                               dp<double>(grams), nkeep);
        else
            hipLaunchKernelGGL(k_gram_mfma, dim3(grid), dim3(256), 0, ctx->stream, nbatch, reps, dp<double>(part),
                               dp<double>(grams), nkeep);
        FEASM_HIP(hipGetLastError());
        FEASM_HIP(hipEventRecord(e1, ctx->stream));
        FEASM_HIP(hipEventSynchronize(e1));
        float t = 0;
        FEASM_HIP(hipEventElapsedTime(&t, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms = t;
        std::vector<double> p((size_t)grid * 256);
        FEASM_HIP(hipMemcpy(p.data(), part->p, p.size() * sizeof(double), hipMemcpyDeviceToHost));
        double s = 0;
        for (double v : p) s += v;
        *checksum = s;
        if (nkeep)
            FEASM_HIP(hipMemcpy(grams_host, grams->p, (size_t)nkeep * 36 * 9 * sizeof(double), hipMemcpyDeviceToHost));
        return FEASM_OK;
    } catch (const feasm::Error &e) {
        if (ctx) ctx->last_error = e.what();
        return e.code;
    }
}
