// Symbolic phase: the device-resident CSC pattern of a mesh's global matrix.
//
// The reference has no symbolic phase: assemble! buffers every triple and
// SparseArrays.sparse discovers the pattern by sorting them on every
// makematrix! (src/AssemblyModule.jl:250-329).  Here the pattern is built
// once per (mesh, ndn) from the connectivity and kept in HBM:
//   1. node -> incident elements, ascending element number (the order in
//      which sparse() meets duplicates: element-major COO);
//   2. node adjacency (sorted, unique) = the node-block pattern; a node pair
//      (m, c) is a block of ndn x ndn entries;
//   3. per block, the list of element contributions (element, upper node
//      pair, transposed?) in ascending element order, so that the numeric
//      phase sums duplicates in exactly the reference's order -- written by
//      the adjacency sort itself when the staged path will run first (keyed
//      candidates), else merged per block when the staged path first needs it;
//   4. the DOF-level CSC colptr/rowval for the node-major numbering
//      dof(node, k) = node*ndn + k (numberdofs! with no EBCs,
//      src/FieldModule.jl:356-377), rows sorted within each column.
#include "common.hpp"

#include <rocprim/rocprim.hpp>

namespace feasm {

namespace {

constexpr int kCap = 256;  // max distinct neighbours of one node

// flags connectivity entries outside [0, nnodes)
__global__ void k_conn_range(const int32_t *__restrict__ conn, int64_t n, int64_t nnodes,
                             unsigned int *__restrict__ bad) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = conn[i];
        if (v < 0 || v >= nnodes) atomicOr(bad, 1u);
    }
}

__global__ void k_n2e_count(const int32_t *__restrict__ conn, int64_t n, int32_t *__restrict__ cnt) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&cnt[conn[i]], 1);
}

// each entry takes a slot of its node's list by counting the node's count
// down (the list is sorted afterwards)
__global__ void k_n2e_fill(const int32_t *__restrict__ conn, int64_t nelem, int npe,
                           const int64_t *__restrict__ ptr, int32_t *__restrict__ cnt, uint32_t *__restrict__ n2e) {
    const int64_t n = nelem * npe;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = i / npe;
        const int a = (int)(i % npe);
        const int32_t c = conn[i];
        const int64_t pos = ptr[c] + (int64_t)atomicSub(&cnt[c], 1) - 1;
        n2e[pos] = ((uint32_t)e << 4) | (uint32_t)a;
    }
}

// Ascending bitonic sort of N (a power of two) values in one thread's
// registers (every index a compile-time constant once unrolled).
template <int N>
__device__ __forceinline__ void thread_sort(uint32_t (&x)[N]) {
#pragma unroll
    for (int k = 2; k <= N; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < N; i++) {
                const int l = i ^ j;
                if (l > i) {
                    const uint32_t a = x[i], b = x[l];
                    const bool up = (i & k) == 0;
                    x[i] = up ? min(a, b) : max(a, b);
                    x[l] = up ? max(a, b) : min(a, b);
                }
            }
}

template <int N>
__device__ __forceinline__ void sort_list(uint32_t *__restrict__ v, int len) {
    uint32_t x[N];
#pragma unroll
    for (int i = 0; i < N; i++) x[i] = i < len ? v[i] : 0xffffffffu;
    thread_sort<N>(x);
#pragma unroll
    for (int i = 0; i < N; i++)
        if (i < len) v[i] = x[i];
}

// the atomic fill leaves each node's list in arbitrary order; sort it (in
// registers up to 32 entries)
__global__ void k_n2e_sort(const int64_t *__restrict__ ptr, uint32_t *__restrict__ n2e, int64_t nnodes) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nnodes;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = ptr[c], e = ptr[c + 1];
        const int len = (int)(e - b);
        if (len <= 8) {
            if (len > 1) sort_list<8>(n2e + b, len);
            continue;
        }
        if (len <= 16) {
            sort_list<16>(n2e + b, len);
            continue;
        }
        if (len <= 32) {
            sort_list<32>(n2e + b, len);
            continue;
        }
        for (int64_t i = b + 1; i < e; i++) {
            const uint32_t v = n2e[i];
            int64_t j = i - 1;
            while (j >= b && n2e[j] > v) {
                n2e[j + 1] = n2e[j];
                j--;
            }
            n2e[j + 1] = v;
        }
    }
}

// Distinct neighbours of node c (including c), sorted: one wavefront (one
// 64-thread block) per node.  The node's candidates (the nodes of its
// elements, npe per element) are gathered into LDS, padded to a power of two
// and sorted with a bitonic network; the distinct values are the ones that
// differ from their predecessor, placed by a prefix count over the lanes'
// runs.  Count pass (adj_ptr == nullptr) writes deg[c]; fill pass writes
// adj[adj_ptr[c] ...].
constexpr int kMaxCand = 2048;  // candidates of one node: incident elements x npe

// Ascending bitonic sort of 64 E values held by one wavefront in
// registers: lane l holds entries l + 64 m.  Partners less than 64 apart
// are exchanged with lane shuffles, partners 64 or more apart sit in two
// registers of the same lane.
template <int E>
__device__ __forceinline__ void wave_sort(int32_t (&x)[E], int lane) {
#pragma unroll
    for (int k = 2; k <= 64 * E; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= 64) {
                const int jj = j >> 6;
#pragma unroll
                for (int m = 0; m < E; m++) {
                    const int mp1 = m ^ 1, mp2 = m ^ 2;
                    const int mp = jj == 1 ? mp1 : mp2;
                    if (mp > m) {
                        const bool up = ((lane + 64 * m) & k) == 0;
                        // both partners named with compile-time indices
                        int32_t &a = x[m];
                        int32_t &b = (jj == 1) ? x[mp1 < E ? mp1 : m] : x[mp2 < E ? mp2 : m];
                        const int32_t lo = min(a, b), hi = max(a, b);
                        a = up ? lo : hi;
                        b = up ? hi : lo;
                    }
                }
            } else {
#pragma unroll
                for (int m = 0; m < E; m++) {
                    const int i = lane + 64 * m;
                    const int32_t y = lane_xor(x[m], j, lane);
                    const bool take_min = ((i & k) == 0) == ((i & j) == 0);
                    x[m] = take_min ? min(x[m], y) : max(x[m], y);
                }
            }
        }
}

// The distinct values v = x >> S of a wave-sorted register list of L keys
// x (the pads after them): their number; each value written to out[pos] and
// its number of occurrences to mult[pos], pos its rank among the distinct
// values.
template <int E, int S>
__device__ __forceinline__ int wave_unique(const int32_t (&x)[E], int lane, int L, int32_t *out, uint16_t *mult) {
    uint64_t fm[E];
    bool first[E];
#pragma unroll
    for (int m = 0; m < E; m++) {
        int32_t prev = __shfl_up(x[m], 1, 64);
        if (m > 0) {
            const int32_t last = __shfl(x[m > 0 ? m - 1 : 0], 63, 64);
            if (lane == 0) prev = last;
        }
        const int i = lane + 64 * m;
        first[m] = i < L && (i == 0 || (x[m] >> S) != (prev >> S));
        fm[m] = __ballot(first[m]);
    }
    int base = 0;
#pragma unroll
    for (int m = 0; m < E; m++) {
        if (first[m]) {
            const int i = lane + 64 * m;
            // the next distinct value's first index (L after the last)
            int nxt = L;
            const uint64_t above = fm[m] & ~((2ull << lane) - 1ull);
            if (above) {
                nxt = 64 * m + __ffsll((unsigned long long)above) - 1;
            } else {
#pragma unroll
                for (int q = E - 1; q > m; q--)
                    if (fm[q]) nxt = 64 * q + __ffsll((unsigned long long)fm[q]) - 1;
            }
            const int pos = base + __popcll(fm[m] & ((1ull << lane) - 1ull));
            out[pos] = x[m] >> S;
            mult[pos] = (uint16_t)(nxt - i);
        }
        base += __popcll(fm[m]);
    }
    return base;
}

// Candidate i of a node (i < L = npe x its elements) is local node i % npe of
// the node's (i / npe)-th element entry.  Keyed (S = 8, L <= 256): the sort
// key is (candidate node << 8) | i, so the candidates of one neighbour come
// out in ascending (element entry, local index) -- the order in which the
// staged path sums the block's contributions -- and, when contrib is given,
// each candidate's contribution record is written at its sorted position
// (the column node's contributions start at npe x its first element entry).
template <int E, int S>
__device__ __forceinline__ int adj_registers(const uint32_t *__restrict__ n2e, const int32_t *__restrict__ conn,
                                             int npe, int64_t s0, int L, int lane, int32_t *out, uint16_t *mult,
                                             uint32_t *__restrict__ contrib) {
    constexpr int32_t kPad = 0x7fffffff;
    int32_t x[E];
#pragma unroll
    for (int m = 0; m < E; m++) {
        const int i = lane + 64 * m;
        x[m] = i < L ? (conn[(int64_t)(n2e[s0 + i / npe] >> 4) * npe + i % npe] << S) | (S ? i : 0) : kPad;
    }
    wave_sort<E>(x, lane);
    const int tot = wave_unique<E, S>(x, lane, L, out, mult);
    if (S && contrib) {
#pragma unroll
        for (int m = 0; m < E; m++) {
            const int i = lane + 64 * m;
            if (i < L) {
                const int ci = x[m] & ((1 << S) - 1);
                const uint32_t u = n2e[s0 + ci / npe];
                const int ac = (int)(u & 15u), am = ci % npe;
                const int lo = am < ac ? am : ac, hi = am < ac ? ac : am;
                const uint32_t p = (uint32_t)(hi * (hi + 1) / 2 + lo);
                contrib[(int64_t)npe * s0 + i] = ((u >> 4) << 7) | (p << 1) | (am > ac ? 1u : 0u);
            }
        }
    }
    return tot;
}

// Node adjacency, one wavefront per node: the nodes of the node's elements,
// sorted, each distinct one with its number of occurrences (one per
// (local index of the node, local index of the neighbour) of every element
// the two share: the block's contributions).  The list of
// node c goes to tmp / tmult at npe * n2e_ptr[c] (it has at most that many
// candidates' room), its length to deg[c]; k_adj_compact packs the lists.
// S = 8 (node numbers below 2^23 - 1) sorts keyed candidates, and with
// contrib also writes the contribution lists; a node with more than 256
// candidates then sets bit 2 of overflow (its lists are left to k_contrib).
template <int S>
__global__ __launch_bounds__(64) void k_adj(const int64_t *__restrict__ n2e_ptr, const uint32_t *__restrict__ n2e,
                                            const int32_t *__restrict__ conn, int npe, int64_t nnodes,
                                            int32_t *__restrict__ deg, int32_t *__restrict__ tmp,
                                            uint16_t *__restrict__ tmult, int32_t *__restrict__ overflow,
                                            uint32_t *__restrict__ contrib) {
    __shared__ int32_t cv[kMaxCand];
    __shared__ int32_t part[64];
    const int lane = threadIdx.x;
    constexpr int32_t kPad = 0x7fffffff;  // above every node number
    for (int64_t c = blockIdx.x; c < nnodes; c += gridDim.x) {
        const int64_t s0 = n2e_ptr[c];
        int L = (int)((n2e_ptr[c + 1] - s0) * npe);
        int32_t *out = tmp + (int64_t)npe * s0;
        uint16_t *mout = tmult + (int64_t)npe * s0;
        if (L <= 256) {  // the candidates fit in the wavefront's registers
            const int tot = L <= 64    ? adj_registers<1, S>(n2e, conn, npe, s0, L, lane, out, mout, contrib)
                            : L <= 128 ? adj_registers<2, S>(n2e, conn, npe, s0, L, lane, out, mout, contrib)
                                       : adj_registers<4, S>(n2e, conn, npe, s0, L, lane, out, mout, contrib);
            if (lane == 0) {
                if (tot > kCap) atomicOr(overflow, 1);
                deg[c] = tot > kCap ? 0 : tot;
            }
            continue;
        }
        if (L > kMaxCand) {
            if (lane == 0) atomicOr(overflow, 1);
            L = 0;
        }
        if (contrib && lane == 0) atomicOr(overflow, 2);
        int P = 1;
        while (P < L) P <<= 1;
        for (int i = lane; i < P; i += 64) {
            int32_t v = kPad;
            if (i < L) v = conn[(int64_t)(n2e[s0 + i / npe] >> 4) * npe + i % npe];
            cv[i] = v;
        }
        __syncthreads();
        for (int k = 2; k <= P; k <<= 1)
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int i = lane; i < P; i += 64) {
                    const int ixj = i ^ j;
                    if (ixj > i) {
                        const int32_t x = cv[i], y = cv[ixj];
                        if ((x > y) == ((i & k) == 0)) {
                            cv[i] = y;
                            cv[ixj] = x;
                        }
                    }
                }
                __syncthreads();
            }
        // each lane takes a run of consecutive sorted entries
        const int per = (P + 63) / 64, beg = min(lane * per, P), end = min(beg + per, P);
        int cnt = 0;
        for (int i = beg; i < end; i++) {
            const int32_t v = cv[i];
            cnt += (v != kPad && (i == 0 || cv[i - 1] != v)) ? 1 : 0;
        }
        part[lane] = cnt;
        __syncthreads();
        int off = 0, tot = 0;
        for (int q = 0; q < 64; q++) {
            const int pq = part[q];
            off += (q < lane) ? pq : 0;
            tot += pq;
        }
        if (tot > kCap) {
            if (lane == 0) {
                atomicOr(overflow, 1);
                deg[c] = 0;
            }
        } else {
            int w = off;
            for (int i = beg; i < end; i++) {
                const int32_t v = cv[i];
                if (v != kPad && (i == 0 || cv[i - 1] != v)) {
                    int r = i + 1;
                    while (r < L && cv[r] == v) r++;
                    out[w] = v;
                    mout[w] = (uint16_t)(r - i);
                    w++;
                }
            }
            if (lane == 0) deg[c] = tot;
        }
        __syncthreads();  // cv and part are refilled for the next node
    }
}

// packs the lists of k_adj: node c's at adj_ptr[c], with the column node of
// each block, and either the multiplicities (bmult) or the contribution list
// offsets (cptr: node c's contributions start at npe * n2e_ptr[c], each
// block's after the ones of the blocks before it in the column)
__global__ __launch_bounds__(64) void k_adj_compact(const int64_t *__restrict__ n2e_ptr, int npe, int64_t nnodes,
                                                    const int64_t *__restrict__ adj_ptr, const int32_t *__restrict__ tmp,
                                                    const uint16_t *__restrict__ tmult, int32_t *__restrict__ adj,
                                                    uint16_t *__restrict__ bmult, int32_t *__restrict__ bcol,
                                                    int64_t *__restrict__ cptr) {
    const int lane = threadIdx.x;
    for (int64_t c = blockIdx.x; c < nnodes; c += gridDim.x) {
        const int64_t src = (int64_t)npe * n2e_ptr[c], dst = adj_ptr[c];
        const int d = (int)(adj_ptr[c + 1] - dst);
        int64_t carry = src;
        for (int i0 = 0; i0 < d; i0 += 64) {  // uniform trip count: the shuffles see every lane
            const int i = i0 + lane;
            const int mlt = i < d ? (int)tmult[src + i] : 0;
            if (i < d) {
                adj[dst + i] = tmp[src + i];
                bcol[dst + i] = (int32_t)c;
                if (bmult) bmult[dst + i] = (uint16_t)mlt;
            }
            if (cptr) {
                int incl = mlt;  // inclusive prefix over the lanes
#pragma unroll
                for (int k = 1; k < 64; k <<= 1) {
                    const int y = __shfl_up(incl, k, 64);
                    if (lane >= k) incl += y;
                }
                if (i < d) cptr[dst + i] = carry + incl - mlt;
                carry += __shfl(incl, 63, 64);
            }
        }
        if (cptr && c == nnodes - 1 && lane == 0) cptr[dst + d] = (int64_t)npe * n2e_ptr[nnodes];
    }
}

// Contributions of block t = (row node m, column node c): the elements
// holding both nodes, ascending -- one thread per block, merging the two
// nodes' element lists (both sorted by element, then local index).  Every
// entry (e, ac) of c's list whose element also holds m gives one record, with
// am the first local index of m in e.  Count pass (cptr == nullptr) or fill
// pass.
__global__ __launch_bounds__(256) void k_contrib(const int32_t *__restrict__ bcol, const int32_t *__restrict__ adj,
                                                 int64_t nblocks, const int64_t *__restrict__ n2e_ptr,
                                                 const uint32_t *__restrict__ n2e, const int64_t *__restrict__ cptr,
                                                 int32_t *__restrict__ ccount, uint32_t *__restrict__ contrib) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nblocks;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c = bcol[t], m = adj[t];
        int cnt = 0;
        int64_t w = cptr ? cptr[t] : 0;
        int64_t j = n2e_ptr[m];
        const int64_t je = n2e_ptr[m + 1];
        // every (local index of c, local index of m) of every shared
        // element: an element that lists a node twice contributes each pair
        for (int64_t s = n2e_ptr[c]; s < n2e_ptr[c + 1]; s++) {
            const uint32_t u = n2e[s];
            const uint32_t e = u >> 4;
            while (j < je && (n2e[j] >> 4) < e) j++;
            if (j == je) break;
            for (int64_t q = j; q < je && (n2e[q] >> 4) == e; q++) {
                if (cptr) {
                    const uint32_t v = n2e[q];
                    const int ac = (int)(u & 15u), am = (int)(v & 15u);
                    const int lo = am < ac ? am : ac, hi = am < ac ? ac : am;
                    const uint32_t p = (uint32_t)(hi * (hi + 1) / 2 + lo);
                    contrib[w++] = (e << 7) | (p << 1) | (am > ac ? 1u : 0u);
                }
                cnt++;
            }
        }
        if (!cptr) ccount[t] = cnt;
    }
}

// One thread per node block t = (row node m, column node c), the k-th block
// of its column: in each of c's NDN DOF columns (NDN deg(c) rows each) the
// block's NDN rows m NDN + j sit at k NDN + j.
template <int NDN>
__global__ __launch_bounds__(256) void k_dof_rows(const int64_t *__restrict__ adj_ptr,
                                                  const int32_t *__restrict__ adj, const int32_t *__restrict__ bcol,
                                                  int64_t nblocks, int32_t *__restrict__ rowval) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nblocks;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c = bcol[t], m = adj[t];
        const int64_t b = adj_ptr[c];
        const int64_t deg = adj_ptr[c + 1] - b, k = t - b;
        const int64_t base = (int64_t)NDN * NDN * b;
#pragma unroll
        for (int i = 0; i < NDN; i++)
#pragma unroll
            for (int j = 0; j < NDN; j++) rowval[base + i * NDN * deg + k * NDN + j] = m * NDN + j;
    }
}

// One thread per node: the column pointers of its NDN DOF columns (a node
// of no element has empty columns), the last thread also the end.
template <int NDN>
__global__ __launch_bounds__(256) void k_dof_colptr(const int64_t *__restrict__ adj_ptr, int64_t nnodes,
                                                    int64_t *__restrict__ colptr) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nnodes;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = adj_ptr[c], deg = adj_ptr[c + 1] - b;
#pragma unroll
        for (int i = 0; i < NDN; i++) colptr[c * NDN + i] = (int64_t)NDN * NDN * b + i * NDN * deg;
        if (c == nnodes - 1) colptr[nnodes * NDN] = (int64_t)NDN * NDN * (b + deg);
    }
}

template <int NDN>
void launch_dof_pattern(feasm_ctx *ctx, const Pattern &P) {
    if (P.nblocks > 0)
        hipLaunchKernelGGL(k_dof_rows<NDN>, dim3(grid_for(P.nblocks, 256)), dim3(256), 0, ctx->stream,
                           dp<int64_t>(P.adj_ptr), dp<int32_t>(P.adj), dp<int32_t>(P.bcol), P.nblocks,
                           dp<int32_t>(P.rowval));
    hipLaunchKernelGGL(k_dof_colptr<NDN>, dim3(grid_for(P.nnodes, 256)), dim3(256), 0, ctx->stream,
                       dp<int64_t>(P.adj_ptr), P.nnodes, dp<int64_t>(P.colptr));
    FEASM_HIP(hipGetLastError());
}

}  // namespace

void exclusive_scan_i32_to_i64(feasm_ctx *ctx, const int32_t *in, int64_t *out, int64_t n) {
    size_t bytes = 0;
    FEASM_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, (int64_t)0, (size_t)n, rocprim::plus<int64_t>(),
                                      ctx->stream));
    void *tmp = ctx->scratch_get(bytes);
    FEASM_HIP(
        rocprim::exclusive_scan(tmp, bytes, in, out, (int64_t)0, (size_t)n, rocprim::plus<int64_t>(), ctx->stream));
}

void exclusive_scan_u16_to_i64(feasm_ctx *ctx, const uint16_t *in, int64_t *out, int64_t n) {
    size_t bytes = 0;
    FEASM_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, (int64_t)0, (size_t)n, rocprim::plus<int64_t>(),
                                      ctx->stream));
    void *tmp = ctx->scratch_get(bytes);
    FEASM_HIP(
        rocprim::exclusive_scan(tmp, bytes, in, out, (int64_t)0, (size_t)n, rocprim::plus<int64_t>(), ctx->stream));
}

void exclusive_scan_i64(feasm_ctx *ctx, const int64_t *in, int64_t *out, int64_t n) {
    size_t bytes = 0;
    FEASM_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, (int64_t)0, (size_t)n, rocprim::plus<int64_t>(),
                                      ctx->stream));
    void *tmp = ctx->scratch_get(bytes);
    FEASM_HIP(
        rocprim::exclusive_scan(tmp, bytes, in, out, (int64_t)0, (size_t)n, rocprim::plus<int64_t>(), ctx->stream));
}

bool conn_in_range(feasm_ctx *ctx, const int32_t *d_conn, int64_t n, int64_t nnodes) {
    if (n == 0) return true;
    DevPtr flag = dev_alloc(sizeof(unsigned int));
    FEASM_HIP(hipMemsetAsync(flag->p, 0, sizeof(unsigned int), ctx->stream));
    hipLaunchKernelGGL(k_conn_range, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, d_conn, n, nnodes,
                       dp<unsigned int>(flag));
    FEASM_HIP(hipGetLastError());
    unsigned int bad = 0;
    read_back(ctx->stream, &bad, flag->p, sizeof(bad));
    return bad == 0;
}

void build_node_to_element(feasm_mesh *m) {
    feasm_ctx *ctx = m->ctx;
    const int64_t nent = m->nelem * m->npe;
    FEASM_REQUIRE(m->nelem < (1ll << 25), FEASM_EUNSUPPORTED, "more than 2^25 elements in one mesh");
    DevPtr cnt = dev_alloc((m->nnodes + 1) * sizeof(int32_t));
    FEASM_HIP(hipMemsetAsync(cnt->p, 0, (m->nnodes + 1) * sizeof(int32_t), ctx->stream));
    hipLaunchKernelGGL(k_n2e_count, dim3(grid_for(nent, 256)), dim3(256), 0, ctx->stream, dp<int32_t>(m->conn),
                       nent, dp<int32_t>(cnt));
    m->n2e_ptr = dev_alloc((m->nnodes + 1) * sizeof(int64_t));
    exclusive_scan_i32_to_i64(ctx, dp<int32_t>(cnt), dp<int64_t>(m->n2e_ptr), m->nnodes + 1);
    m->n2e = dev_alloc(nent * sizeof(uint32_t));
    hipLaunchKernelGGL(k_n2e_fill, dim3(grid_for(nent, 256)), dim3(256), 0, ctx->stream, dp<int32_t>(m->conn),
                       m->nelem, m->npe, dp<int64_t>(m->n2e_ptr), dp<int32_t>(cnt), dp<uint32_t>(m->n2e));
    hipLaunchKernelGGL(k_n2e_sort, dim3(grid_for(m->nnodes, 256)), dim3(256), 0, ctx->stream,
                       dp<int64_t>(m->n2e_ptr), dp<uint32_t>(m->n2e), m->nnodes);
    FEASM_HIP(hipGetLastError());
    // no wait: the arrays are read by later work on the same stream (the
    // temporaries above are released behind a fence on it)
}

void build_contributions(feasm_mesh *m, Pattern &P_) {
    feasm_ctx *ctx = m->ctx;
    Pattern *P = &P_;
    if (P->contrib && P->cptr) return;
    if (P->cptr) {
        // offsets written with the pattern
    } else if (P->bmult) {
        P->cptr = dev_alloc((P->nblocks + 1) * sizeof(int64_t));
        // a block's contributions are its row node's occurrences among the
        // column node's elements' nodes, counted by the adjacency pass
        exclusive_scan_u16_to_i64(ctx, dp<uint16_t>(P->bmult), dp<int64_t>(P->cptr), P->nblocks + 1);
    } else {
        P->cptr = dev_alloc((P->nblocks + 1) * sizeof(int64_t));
        DevPtr ccount = dev_alloc((P->nblocks + 1) * sizeof(int32_t));
        FEASM_HIP(hipMemsetAsync(ccount->p, 0, (P->nblocks + 1) * sizeof(int32_t), ctx->stream));
        hipLaunchKernelGGL(k_contrib, dim3(grid_for(P->nblocks, 256)), dim3(256), 0, ctx->stream,
                           dp<int32_t>(P->bcol), dp<int32_t>(P->adj), P->nblocks, dp<int64_t>(m->n2e_ptr),
                           dp<uint32_t>(m->n2e), (const int64_t *)nullptr, dp<int32_t>(ccount), (uint32_t *)nullptr);
        FEASM_HIP(hipGetLastError());
        exclusive_scan_i32_to_i64(ctx, dp<int32_t>(ccount), dp<int64_t>(P->cptr), P->nblocks + 1);
    }
    // every (element, local column, local row) is one contribution
    P->ncontrib = m->nelem * m->npe * m->npe;
    if (P->contrib) {  // written by the adjacency pass
        P->bmult.reset();
        return;
    }
    DevPtr contrib = dev_alloc(std::max<size_t>(P->ncontrib * sizeof(uint32_t), 4));
    hipLaunchKernelGGL(k_contrib, dim3(grid_for(P->nblocks, 256)), dim3(256), 0, ctx->stream, dp<int32_t>(P->bcol),
                       dp<int32_t>(P->adj), P->nblocks, dp<int64_t>(m->n2e_ptr), dp<uint32_t>(m->n2e),
                       dp<int64_t>(P->cptr), (int32_t *)nullptr, dp<uint32_t>(contrib));
    FEASM_HIP(hipGetLastError());
    P->contrib = contrib;
    P->bmult.reset();
}

std::shared_ptr<Pattern> build_pattern(feasm_mesh *m, int ndn, bool with_contrib) {
    feasm_ctx *ctx = m->ctx;
    if (!m->n2e_ptr) build_node_to_element(m);
    auto P = std::make_shared<Pattern>();
    P->ndn = ndn;
    P->nnodes = m->nnodes;
    const int64_t nn = m->nnodes;
    // 2. adjacency: one pass into per-node room, then packed
    const int npe = m->npe;
    const int64_t nent = m->nelem * npe;  // node -> element entries
    // degrees [nn + 1] (the last one zero, for the scan) and the flags word
    // after them (too many neighbours; lists left to k_contrib): one clear
    DevPtr deg = dev_alloc((nn + 2) * sizeof(int32_t));
    int32_t *flags = dp<int32_t>(deg) + nn + 1;
    DevPtr tmp = dev_alloc(std::max<size_t>((size_t)npe * nent * sizeof(int32_t), 4));
    DevPtr tmult = dev_alloc(std::max<size_t>((size_t)npe * nent * sizeof(uint16_t), 4));
    FEASM_HIP(hipMemsetAsync(dp<int32_t>(deg) + nn, 0, 2 * sizeof(int32_t), ctx->stream));
    const unsigned gadj = grid_for(nn, 1, 1 << 17);
    // the staged path's contribution lists come out of the adjacency sort
    // when the node numbers leave room for the candidate index in the keys
    const bool keyed = with_contrib && nn < (1ll << 23) - 1;
    if (keyed) {
        P->ncontrib = m->nelem * npe * npe;
        P->contrib = dev_alloc(std::max<size_t>(P->ncontrib * sizeof(uint32_t), 4));
    }
    uint32_t *cout = keyed ? dp<uint32_t>(P->contrib) : nullptr;
    if (keyed)
        hipLaunchKernelGGL(k_adj<8>, dim3(gadj), dim3(64), 0, ctx->stream, dp<int64_t>(m->n2e_ptr),
                           dp<uint32_t>(m->n2e), dp<int32_t>(m->conn), npe, nn, dp<int32_t>(deg), dp<int32_t>(tmp),
                           dp<uint16_t>(tmult), flags, cout);
    else
        hipLaunchKernelGGL(k_adj<0>, dim3(gadj), dim3(64), 0, ctx->stream, dp<int64_t>(m->n2e_ptr),
                           dp<uint32_t>(m->n2e), dp<int32_t>(m->conn), npe, nn, dp<int32_t>(deg), dp<int32_t>(tmp),
                           dp<uint16_t>(tmult), flags, cout);
    FEASM_HIP(hipGetLastError());
    P->adj_ptr = dev_alloc((nn + 1) * sizeof(int64_t));
    exclusive_scan_i32_to_i64(ctx, dp<int32_t>(deg), dp<int64_t>(P->adj_ptr), nn + 1);
    int32_t fl = 0;
    const ReadBack rb[2] = {{&P->nblocks, dp<int64_t>(P->adj_ptr) + nn, sizeof(int64_t)},
                            {&fl, flags, sizeof(int32_t)}};
    read_back(ctx->stream, rb, 2);
    FEASM_REQUIRE(!(fl & 1), FEASM_EUNSUPPORTED,
                  "a node has more than 256 distinct neighbours or more than 2048 (element, node) incidences");
    if (fl & 2) P->contrib.reset();  // some node's lists are left to k_contrib
    P->adj = dev_alloc(std::max<size_t>(P->nblocks * sizeof(int32_t), 4));
    if (with_contrib) {
        // the staged path's list offsets, straight from the multiplicities
        P->cptr = dev_alloc((P->nblocks + 1) * sizeof(int64_t));
        if (nn == 0) FEASM_HIP(hipMemsetAsync(P->cptr->p, 0, sizeof(int64_t), ctx->stream));
    } else {
        P->bmult = dev_alloc((P->nblocks + 1) * sizeof(uint16_t));
        FEASM_HIP(hipMemsetAsync(dp<uint16_t>(P->bmult) + P->nblocks, 0, sizeof(uint16_t), ctx->stream));
    }
    P->bcol = dev_alloc(std::max<size_t>(P->nblocks * sizeof(int32_t), 4));
    hipLaunchKernelGGL(k_adj_compact, dim3(gadj), dim3(64), 0, ctx->stream, dp<int64_t>(m->n2e_ptr), npe, nn,
                       dp<int64_t>(P->adj_ptr), dp<int32_t>(tmp), dp<uint16_t>(tmult), dp<int32_t>(P->adj),
                       P->bmult ? dp<uint16_t>(P->bmult) : nullptr, dp<int32_t>(P->bcol),
                       P->cptr ? dp<int64_t>(P->cptr) : nullptr);
    FEASM_HIP(hipGetLastError());
    // 4. DOF-level CSC
    FEASM_REQUIRE(ndn >= 1 && ndn <= 3, FEASM_EUNSUPPORTED, "1 to 3 DOFs per node");
    P->n = nn * ndn;
    P->nnz = P->nblocks * ndn * ndn;
    FEASM_REQUIRE(P->n < (1ll << 31), FEASM_EUNSUPPORTED, "more than 2^31 DOFs");
    P->colptr = dev_alloc((P->n + 1) * sizeof(int64_t));
    P->rowval = dev_alloc(std::max<size_t>(P->nnz * sizeof(int32_t), 4));
    if (nn == 0) {
        FEASM_HIP(hipMemsetAsync(P->colptr->p, 0, sizeof(int64_t), ctx->stream));
    } else if (ndn == 1) {
        launch_dof_pattern<1>(ctx, *P);
    } else if (ndn == 2) {
        launch_dof_pattern<2>(ctx, *P);
    } else {
        launch_dof_pattern<3>(ctx, *P);
    }
    return P;
}

}  // namespace feasm
