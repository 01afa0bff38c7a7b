// feasm_sparse_device_noalloc: the device sparse() of coo.hip (same kernels,
// same additions in the same order) into caller-provided outputs, with every
// temporary in one fixed device array instead of allocated memory.
#include <mutex>

#include "common.hpp"

namespace feasm {

namespace {

constexpr int64_t kNoAllocMaxTriples = FEASM_NOALLOC_MAX_TRIPLES;
constexpr size_t kNoAllocTmpBytes = (size_t)4 << 20;  // sort and scan storage
// per triple: sort keys (8 B) and their ping-pong copy (8 B), segment starts
// (8 B), sort payload (4 B) and its copy (4 B); then 256 B of counters
constexpr size_t kNoAllocBytes = (size_t)kNoAllocMaxTriples * (8 + 8 + 8 + 4 + 4) + 256 + kNoAllocTmpBytes;

// Every temporary of feasm_sparse_device_noalloc lives here.  Each GPU holds
// its own copy, at its own address, until the process ends; every call on a
// GPU reuses it and keeps the last call's contents, so calls that overlap on
// one GPU race.  Host code never passes this name to a kernel (on the host it
// is only a stand-in): it uses the device address looked up below.
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// device address of this GPU's copy of `scratch`, looked up on the first call
// on the device with that device current, and kept
void *scratch_address(int device) {
    FEASM_REQUIRE(device >= 0 && device < kMaxDevices, FEASM_EUNSUPPORTED, "device index beyond 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        FEASM_HIP(hipGetDevice(&cur));
        if (cur != device) FEASM_HIP(hipSetDevice(device));
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        FEASM_HIP(e);
        FEASM_REQUIRE(ptr != nullptr, FEASM_EHIP, "no device address for the scratch array");
        g_scratch_ptr[device] = ptr;
    }
    return g_scratch_ptr[device];
}

thread_local std::string g_noalloc_error;

}  // namespace

}  // namespace feasm

using namespace feasm;

extern "C" int feasm_sparse_device_noalloc(feasm_ctx *ctx, int64_t ntrip, const int32_t *d_I, const int32_t *d_J,
                                           const double *d_V, int64_t m, int64_t n, int64_t *d_colptr,
                                           int32_t *d_rowval, double *d_nzval, int64_t *nnz) {
    try {
        FEASM_REQUIRE(ctx && d_colptr && nnz, FEASM_EINVAL, "null argument");
        FEASM_REQUIRE(ntrip == 0 || (d_I && d_J && d_V && d_rowval && d_nzval), FEASM_EINVAL, "null argument");
        check_sparse_args(ntrip, m, n);
        FEASM_REQUIRE(ntrip <= kNoAllocMaxTriples, FEASM_ENOMEM,
                      "feasm_sparse_device_noalloc: more than FEASM_NOALLOC_MAX_TRIPLES (1048576) triples");
        hipStream_t s = ctx->stream;
        unsigned char *base = static_cast<unsigned char *>(scratch_address(ctx->device));
        if (ntrip == 0) {
            fill_i64(s, d_colptr, n + 1, 0);
            FEASM_HIP(hipStreamSynchronize(s));
            *nnz = 0;
            return FEASM_OK;
        }
        const unsigned end_bit = (unsigned)bits_for((uint64_t)m * (uint64_t)n);
        const size_t T = (size_t)kNoAllocMaxTriples;
        SparseWork w;
        w.keys = reinterpret_cast<uint64_t *>(base);
        w.keys_alt = reinterpret_cast<uint64_t *>(base + 8 * T);
        w.start = reinterpret_cast<int64_t *>(base + 16 * T);
        w.idx = reinterpret_cast<uint32_t *>(base + 24 * T);
        w.idx_alt = reinterpret_cast<uint32_t *>(base + 28 * T);
        w.count = reinterpret_cast<int64_t *>(base + 32 * T);
        w.flag = reinterpret_cast<unsigned int *>(base + 32 * T + 16);
        w.tmp = base + 32 * T + 256;
        w.tmp_bytes = sort_tmp_bytes(ntrip, end_bit, s);
        FEASM_REQUIRE(w.tmp_bytes <= kNoAllocTmpBytes, FEASM_ENOMEM,
                      "feasm_sparse_device_noalloc: sort storage beyond the fixed scratch array");
        w.tmp_bytes = kNoAllocTmpBytes;
        const int64_t count = sparse_phase_a(s, ntrip, d_I, d_J, m, n, end_bit, w);
        sparse_phase_b(s, ntrip, d_V, m, n, count, w, d_colptr, d_rowval, d_nzval);
        FEASM_HIP(hipStreamSynchronize(s));
        *nnz = count;
        return FEASM_OK;
    } catch (const feasm::Error &e) {
        g_noalloc_error = e.what();
        if (ctx) ctx->last_error = e.what();
        return e.code;
    } catch (const std::exception &e) {
        g_noalloc_error = e.what();
        if (ctx) ctx->last_error = e.what();
        return FEASM_EINVAL;
    }
}
