// Internal declarations shared by the translation units of libfeasm.so.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/feasm.h"

namespace feasm {

// ---------------------------------------------------------------- errors
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define FEASM_HIP(call)                                                                              \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            throw ::feasm::Error(e_ == hipErrorOutOfMemory ? FEASM_ENOMEM : FEASM_EHIP,              \
                                 std::string("HIP error in ") + #call + ": " + hipGetErrorString(e_)); \
    } while (0)

#define FEASM_REQUIRE(cond, code, msg)                         \
    do {                                                       \
        if (!(cond)) throw ::feasm::Error((code), (msg));      \
    } while (0)

// ---------------------------------------------------------------- device buffers
// pool.cpp: cached device blocks (bytes receives the capacity, >= n)
void *pool_alloc(size_t n, size_t *bytes, int *device);
void pool_release(void *p, size_t bytes, int device);
void pool_register_stream(int device, hipStream_t s);
void pool_unregister_stream(int device, hipStream_t s);
size_t pool_cached_bytes(int device);
// frees every cached block after a device synchronisation; the bytes freed
size_t pool_trim(int device);
// the calling thread's buffers are used on stream s only (null: any
// stream): blocks it releases are fenced on s alone
void pool_set_thread_fence(int device, hipStream_t s);
// pool_set_thread_fence for a scope (the previous setting comes back)
struct ScopedThreadFence {
    hipStream_t prev_s = nullptr;
    int prev_d = -1;
    ScopedThreadFence(int device, hipStream_t s);
    ~ScopedThreadFence();
};
// copies bytes from device memory after the work queued on s, through a
// pinned buffer of the calling thread (a copy into pageable memory goes
// through the runtime's staging path, tens of microseconds a call)
void read_back(hipStream_t s, void *host, const void *dev, size_t bytes);
// several pieces behind one wait
struct ReadBack {
    void *host;
    const void *dev;
    size_t bytes;
};
void read_back(hipStream_t s, const ReadBack *parts, int n);

// Owning device allocation; shared so that a matrix can keep the pattern
// arrays of its mesh alive.  bytes is the capacity (at least what was asked).
struct DevMem {
    void *p = nullptr;
    size_t bytes = 0;
    int device = 0;
    DevMem() = default;
    explicit DevMem(size_t n) {
        if (n) p = pool_alloc(n, &bytes, &device);
    }
    ~DevMem() {
        if (p) pool_release(p, bytes, device);
    }
    DevMem(const DevMem &) = delete;
    DevMem &operator=(const DevMem &) = delete;
};
using DevPtr = std::shared_ptr<DevMem>;

inline DevPtr dev_alloc(size_t bytes) { return std::make_shared<DevMem>(bytes); }

template <typename T>
inline T *dp(const DevPtr &d) {
    return d ? static_cast<T *>(d->p) : nullptr;
}

// ---------------------------------------------------------------- quadrature tables
// Basis values and parametric gradients at the quadrature points of one
// (element type, rule) pair, as integrationdata() returns them
// (src/IntegDomainModule.jl:632-648).  Layout: N[q*npe + a],
// dN[(q*npe + a)*3 + k], w[q].
struct QTable {
    int etype = 0, rule = 0, order = 0, npe = 0, nqp = 0;
    std::vector<double> N, dN, w;
};
QTable make_qtable(int etype, int rule, int order);  // tables.cpp

// Device copy of a QTable, one contiguous buffer: [N | dN | w].
struct DevQTable {
    int npe = 0, nqp = 0;
    DevPtr buf;
    const double *N() const { return dp<double>(buf); }
    const double *dN() const { return dp<double>(buf) + npe * nqp; }
    const double *w() const { return dp<double>(buf) + 4 * npe * nqp; }
};

}  // namespace feasm

namespace feasm {
struct TileSet;
}

// ---------------------------------------------------------------- handles
struct feasm_ctx {
    // handles made on a context (meshes, matrices, assemblers, batches) keep
    // it alive: feasm_ctx_destroy drops the caller's reference and the
    // context is freed with the last one
    std::atomic<int> refs{1};
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    // timing ring: when on, every numeric assembly records three events
    // (before the element kernel, between the kernels, after the gather) in
    // the next slot of the ring
    bool timing = false;
    int ring_slots = 0, ring_next = 0, ring_used = 0;
    std::vector<hipEvent_t> ring;  // 3 events per slot
    void timing_configure(int slots);
    hipEvent_t ring_event(int k) { return ring[3 * ring_next + k]; }
    void ring_advance() {
        ring_next = (ring_next + 1) % ring_slots;
        if (ring_used < ring_slots) ring_used++;
    }
    std::map<int, feasm::DevQTable> qtables;  // key = etype*1000 + rule*100 + order
    // reusable scratch for library temporaries (sort/scan storage)
    feasm::DevPtr scratch;
    void *scratch_get(size_t bytes);
};

// releases the context when the last reference goes (capi.cpp)
void feasm_ctx_release(feasm_ctx *ctx);

// counted reference to a context
struct CtxRef {
    feasm_ctx *p = nullptr;
    CtxRef() = default;
    CtxRef(feasm_ctx *c) : p(c) {
        if (p) p->refs++;
    }
    CtxRef(const CtxRef &o) : CtxRef(o.p) {}
    CtxRef &operator=(feasm_ctx *c) {
        if (c) c->refs++;
        feasm_ctx *old = p;
        p = c;
        if (old) feasm_ctx_release(old);
        return *this;
    }
    CtxRef &operator=(const CtxRef &o) { return (*this = o.p); }
    ~CtxRef() {
        if (p) feasm_ctx_release(p);
    }
    operator feasm_ctx *() const { return p; }
    feasm_ctx *operator->() const { return p; }
};

struct Pattern {
    int ndn = 0;
    int64_t nnodes = 0, nblocks = 0, ncontrib = 0, n = 0, nnz = 0;
    feasm::DevPtr adj_ptr;   // int64 [nnodes+1] node-block row pointer per column node
    feasm::DevPtr adj;       // int32 [nblocks]  row node of each block (sorted per column node)
    feasm::DevPtr bcol;      // int32 [nblocks]  column node of each block
    feasm::DevPtr cptr;      // int64 [nblocks+1] contribution list pointer
    feasm::DevPtr contrib;   // uint32 [ncontrib] (e << 7) | (pair << 1) | transposed
    feasm::DevPtr colptr;    // int64 [n+1]  DOF-level CSC column pointer (0-based)
    feasm::DevPtr rowval;    // int32 [nnz]  DOF-level CSC row index (0-based)
    feasm::DevPtr bmult;     // uint16 [nblocks+1] contributions of each block (until the lists are built)
};

struct feasm_mesh {
    CtxRef ctx;
    int etype = 0, npe = 0;
    int64_t nelem = 0, nnodes = 0;
    feasm::DevPtr conn;     // int32 [nelem*npe] 0-based
    feasm::DevPtr xyz;      // double [nnodes*3]
    feasm::DevPtr n2e_ptr;  // int64 [nnodes+1]
    feasm::DevPtr n2e;      // uint32 [nelem*npe]  (e << 4) | local index, ascending e
    int ndn = 0;
    int64_t nalldofs = 0;
    bool nodemajor = false;
    feasm::DevPtr dofs;     // int32 [nnodes*ndn] 0-based DOF numbers (node-major rows)
    std::shared_ptr<Pattern> pattern;
    feasm::DevPtr stage;    // per-element staging of the node-pair blocks
    // numeric path: 0 = automatic (fused tiles when the element/rule has a
    // fused kernel), 1 = staged element blocks + ordered gather, 2 = fused
    int numeric_path = 0;
    int tile_nodes = 0, tile_kib = 0, tile_threads = 0;  // fused-kernel tile parameters (0 = default)
    int64_t assemblies = 0;  // numeric assemblies run by the assembly server on this mesh
    bool dofs_unique = true;  // no DOF number is shared by two (node, component) pairs
    std::map<int64_t, std::shared_ptr<feasm::TileSet>> tiles;  // keyed by TileShape::key()
};

struct feasm_matrix {
    CtxRef ctx;
    int64_t m = 0, n = 0, nnz = 0;
    feasm::DevPtr colptr;  // int64 [n+1]
    feasm::DevPtr rowval;  // int32 [nnz]
    feasm::DevPtr nzval;   // double [nnz]
    std::shared_ptr<Pattern> pattern;  // set when the values sit in a mesh's resident pattern
    int form_class = -1;               // ndn of the blocks (for bilform_into checks)
};

namespace feasm {

// The value of lane (lane ^ j) of a wavefront, j a power of two below 64:
// exchanges within rows of 16 lanes as DPP moves (quad permutations, row
// rotations: a rotation by r gives lane i the value of lane i - r of its
// row), across 16 lanes as a swizzle, across the halves as a permute.
__device__ __forceinline__ int32_t lane_xor(int32_t x, int j, int lane) {
    switch (j) {
        case 1: return __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
        case 2: return __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
        case 4: {
            const int32_t from_below = __builtin_amdgcn_mov_dpp(x, 0x124, 0xF, 0xF, false);  // row_ror 4: lane i - 4
            const int32_t from_above = __builtin_amdgcn_mov_dpp(x, 0x12C, 0xF, 0xF, false);  // row_ror 12: lane i + 4
            return (lane & 4) ? from_below : from_above;
        }
        case 8: return __builtin_amdgcn_mov_dpp(x, 0x128, 0xF, 0xF, false);  // row_ror 8
        case 16: return __builtin_amdgcn_ds_swizzle(x, 0x401F);              // xor 16 within 32 lanes
        default: return __shfl_xor(x, j, 64);
    }
}

// fused.hip: D (6 x 6, row-major) has cubic symmetry -- D00 = D11 = D22,
// D01 = D02 = D12, D33 = D44 = D55, every other entry zero
bool cubic_symmetry(const double *D);

// tables.cpp
const DevQTable &device_qtable(feasm_ctx *ctx, int etype, int rule, int order);

// vector.hip: load vector of a constant force (ndn values) into F (nalldofs,
// device); stage holds nelem*npe*ndn doubles
void launch_linform_dot(feasm_ctx *ctx, const feasm_mesh *m, const DevQTable &qt, const double *force,
                        double *stage, double *F);

// lumped.hip: HRZ-lumped diagonal of the DOT form with an ndn x ndn
// coefficient (host); a new diagonal CSC matrix
feasm_matrix *hrz_lumped(feasm_ctx *ctx, feasm_mesh *m, const DevQTable &qt, const double *coef_host);

// dropzeros.hip: a new matrix without the stored exact zeros of a
feasm_matrix *drop_zeros(feasm_ctx *ctx, const feasm_matrix *a);

// blocks.hip: principal submatrix on the kept DOFs (host flags, one per
// column), renumbered consecutively
feasm_matrix *principal_submatrix(feasm_ctx *ctx, const feasm_matrix *a, const uint8_t *keep_host);

// blocks.hip: A[1:r, 1:c] (matrix_blocked_ff)
feasm_matrix *leading_block(feasm_ctx *ctx, const feasm_matrix *a, int64_t r, int64_t c);

// symm.hip: S + transpose(S) with the diagonal halved and exact zeros dropped
feasm_matrix *symm_plus_transpose(feasm_ctx *ctx, const feasm_matrix *S);

// pattern.hip
void build_node_to_element(feasm_mesh *m);
// with_contrib: the staged path will run first, so the adjacency pass also
// writes the contribution lists (build_contributions then only scans)
std::shared_ptr<Pattern> build_pattern(feasm_mesh *m, int ndn, bool with_contrib = false);
// every entry of the device connectivity d_conn[0..n) lies in [0, nnodes)
bool conn_in_range(feasm_ctx *ctx, const int32_t *d_conn, int64_t n, int64_t nnodes);
void build_contributions(feasm_mesh *m, Pattern &p);

// element.hip
struct FormParams {
    int form = 0;
    int ndn = 1;
    double coef[36];  // iso: c; mat: kappa 3x3; elastic: D 6x6 (row-major)
};
// number of upper-triangle node pairs and values per pair in the staging
int npairs(int npe);
int stage_width(int form);  // 9 for elasticity/matrix diffusion... see element.hip
void launch_element_stage(feasm_ctx *ctx, const feasm_mesh *m, const DevQTable &qt, const FormParams &fp,
                          int precision, void *stage);
void launch_element_coo(feasm_ctx *ctx, const feasm_mesh *m, const DevQTable &qt, const FormParams &fp,
                        int precision, int32_t *I, int32_t *J, double *V);
void launch_gather(feasm_ctx *ctx, const feasm_mesh *m, const Pattern &p, const FormParams &fp, int precision,
                   const void *stage, double *nzval);

// coo.hip
feasm_matrix *sparse_device(feasm_ctx *ctx, int64_t ntrip, const int32_t *I, const int32_t *J, const double *V,
                            int64_t m, int64_t n);

// scan helpers (coo.hip)
void exclusive_scan_i64(feasm_ctx *ctx, const int64_t *in, int64_t *out, int64_t n);
void exclusive_scan_u16_to_i64(feasm_ctx *ctx, const uint16_t *in, int64_t *out, int64_t n);
void exclusive_scan_i32_to_i64(feasm_ctx *ctx, const int32_t *in, int64_t *out, int64_t n);

inline unsigned grid_for(int64_t n, int block, unsigned cap = 1u << 20) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > (int64_t)cap) g = cap;
    return (unsigned)g;
}

}  // namespace feasm

namespace feasm {

// coo.hip: the two phases of the device sparse().  Temporaries of one call
// on nt triples: the sort ping-pongs between (keys, keys_alt) and (idx,
// idx_alt); afterwards the buffers not holding the sorted data take the
// segment heads (4 B per triple) and numbers (8 B).
struct SparseWork {
    uint64_t *keys = nullptr, *keys_alt = nullptr;
    uint32_t *idx = nullptr, *idx_alt = nullptr;
    int64_t *start = nullptr;  // nnz entries (phase B)
    void *tmp = nullptr;       // sort and scan storage
    size_t tmp_bytes = 0;
    unsigned int *flag = nullptr;
    int64_t *count = nullptr;  // [nnz, flag] read back after phase A
    // where phase A left the sorted keys / payload and the segment arrays
    const uint64_t *skeys = nullptr;
    const uint32_t *sidx = nullptr;
    int32_t *head = nullptr;
    int64_t *seg = nullptr;
};
int bits_for(uint64_t v);
// sort + scan storage for nt triples over end_bit key bits
size_t sort_tmp_bytes(int64_t nt, unsigned end_bit, hipStream_t s);
// keys + range check, stable sort, segment heads and numbers, one read-back:
// returns nnz, throws FEASM_EDOF for an index outside [0,m) x [0,n)
int64_t sparse_phase_a(hipStream_t s, int64_t nt, const int32_t *I, const int32_t *J, int64_t m, int64_t n,
                       unsigned end_bit, SparseWork &w);
// segment starts, ordered sums, column pointers
void sparse_phase_b(hipStream_t s, int64_t nt, const double *V, int64_t m, int64_t n, int64_t nnz,
                    const SparseWork &w, int64_t *colptr, int32_t *rowval, double *nzval);
void check_sparse_args(int64_t nt, int64_t m, int64_t n);
void fill_i64(hipStream_t s, int64_t *p, int64_t n, int64_t v);

}  // namespace feasm
