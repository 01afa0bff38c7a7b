// Tiles for the fused numeric kernel, built on the device.
//
// The same tiles as the host builder (tiles.cpp), byte for byte, from the
// mesh's device arrays, on the mesh's context stream:
//  1. Morton codes of the nodes (fused.hpp: morton_code; the bounding box is
//     reduced on the device, the lattice step computed on the host by the
//     function the host builder uses) and a stable radix sort of
//     (code, node): the host builder's node order.
//  2. The box tree (fused.hpp: box_fits, box_split), one level per launch:
//     one workgroup per box either splits it or evaluates it -- the elements
//     of its nodes and then the nodes of those elements are sorted in LDS and
//     counted -- and marks a tile where it fits.  Every box of a level is
//     disjoint from the others, so a level holds at most nn boxes.
//  3. One workgroup per tile does what the host's make_tile does: the sorted
//     halo and halo-node lists and the halo connectivity; one block task per
//     node block (m, c) of every owned column, its elements found by merging
//     the sorted element lists of the two nodes (ascending element, then
//     local index: the order of the host's column walk); blocks whose row
//     node is owned by the tile folded into their transposes; the stable sort
//     by (element count, symmetric diagonal); the wavefront schedule; the
//     items; for H8 in FP64 the bank-spread halo numbering -- and writes the
//     blob into a slot sized from the partition's upper bounds.
//  4. The blobs are packed (a scan of their sizes) and the headers written.
#include "common.hpp"

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "fused.hpp"

namespace feasm {

namespace {

constexpr int kTB = 256;                 // threads of the box and tile workgroups
constexpr int kMaxNpe = 10;              // element types with a fused kernel
constexpr int kMaxHaloNodes = 255 * kMaxNpe;

struct MeshView {
    const int32_t *conn;
    const int64_t *n2e_ptr;
    const uint32_t *n2e;
    const int64_t *adj_ptr;
    const int32_t *adj;
    int npe;
};

// error bits raised by the kernels
enum : unsigned {
    kErrNodeTooBig = 1u,
    kErrQueue = 2u,
    kErrSizes = 4u,
    kErrDegree = 8u,
    kErrShared = 16u,
    kErrEmptyBlock = 32u,
    kErrAsym = 64u,
    kErrItems = 128u,
    kErrSets = 256u,
    kErrBudget = 512u,
};

// LDS of one box evaluation (sorting buffer, halo lists, owned nodes)
struct EvalLds {
    uint32_t buf[kMaxBoxCand];
    int32_t hn[kMaxHaloNodes + 2];
    int32_t halo[256];
    int32_t owned[256];
    uint32_t coff[260];
    uint32_t part[kTB + 1];
    int sk;
    unsigned long long snt;
};

// ---- workgroup primitives (blockDim.x == kTB)

// exclusive scan of part[0, kTB) in place; part[kTB] = total
__device__ void scan_parts(uint32_t *part) {
    const int t = threadIdx.x;
    __syncthreads();
    if (t < 64) {
        uint32_t v[4], tot = 0;
        for (int i = 0; i < 4; i++) {
            v[i] = part[4 * t + i];
            tot += v[i];
        }
        uint32_t inc = tot;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(inc, d, 64);
            if (t >= d) inc += y;
        }
        uint32_t acc = inc - tot;
        for (int i = 0; i < 4; i++) {
            part[4 * t + i] = acc;
            acc += v[i];
        }
        if (t == 63) part[kTB] = inc;
    }
    __syncthreads();
}

// exclusive scan of a[0, n) in place; returns the total
__device__ uint32_t block_excl_scan(uint32_t *a, int n, uint32_t *part) {
    const int t = threadIdx.x;
    const int per = (n + kTB - 1) / kTB;
    const int b = min(n, t * per), e = min(n, b + per);
    uint32_t s = 0;
    for (int i = b; i < e; i++) s += a[i];
    part[t] = s;
    scan_parts(part);
    uint32_t acc = part[t];
    for (int i = b; i < e; i++) {
        const uint32_t x = a[i];
        a[i] = acc;
        acc += x;
    }
    const uint32_t total = part[kTB];
    __syncthreads();
    return total;
}

// Ascending bitonic sort of a[0, n), n a power of two, 2 <= n <= 4096.
// Up to 1024 entries the entries stay in registers: thread t holds the entries t + kTB m
// (m < E = n / kTB; n below kTB is padded with 0xffffffff to kTB).  A stage
// whose partners are 256 or more apart pairs two registers of one thread,
// one whose partners are less than 64 apart pairs the same register of two
// lanes of one wavefront (a lane shuffle), and the stages in between go
// through LDS with workgroup barriers.
template <int E, int JJ>
__device__ __forceinline__ void sort_reg_stage(uint32_t (&x)[E], int t, int k) {
#pragma unroll
    for (int m = 0; m < E; m++) {
        const int mp = m ^ JJ;
        if (mp > m && mp < E) {
            const bool up = ((t + kTB * m) & k) == 0;
            const uint32_t lo = min(x[m], x[mp]), hi = max(x[m], x[mp]);
            x[m] = up ? lo : hi;
            x[mp] = up ? hi : lo;
        }
    }
}

template <int E>
__device__ void block_sort_e(uint32_t *a) {
    // n = E kTB entries: every stage's distance is a compile-time constant
    // once the loops are unrolled, so the lane exchanges are DPP moves
    constexpr int N = E * kTB;
    const int t = threadIdx.x;
    const int lane = t & 63;
    uint32_t x[E];
#pragma unroll
    for (int m = 0; m < E; m++) x[m] = a[t + kTB * m];
#pragma unroll
    for (int k = 2; k <= N; k <<= 1) {
        int j = k >> 1;
#pragma unroll
        for (; j >= kTB; j >>= 1) {
            const int jj = j / kTB;
            if (jj == 1) sort_reg_stage<E, 1>(x, t, k);
            if (jj == 2) sort_reg_stage<E, 2>(x, t, k);
            if (jj == 4) sort_reg_stage<E, 4>(x, t, k);
            if (jj == 8) sort_reg_stage<E, 8>(x, t, k);
        }
        if (j >= 64) {
            __syncthreads();
#pragma unroll
            for (int m = 0; m < E; m++) a[t + kTB * m] = x[m];
            __syncthreads();
#pragma unroll
            for (; j >= 64; j >>= 1) {
#pragma unroll
                for (int m = 0; m < E; m++) {
                    const int i = t + kTB * m, ixj = i ^ j;
                    if (ixj > i) {
                        const uint32_t p = a[i], q = a[ixj];
                        if ((p > q) == ((i & k) == 0)) {
                            a[i] = q;
                            a[ixj] = p;
                        }
                    }
                }
                __syncthreads();
            }
#pragma unroll
            for (int m = 0; m < E; m++) x[m] = a[t + kTB * m];
        }
#pragma unroll
        for (; j > 0; j >>= 1) {
#pragma unroll
            for (int m = 0; m < E; m++) {
                const int i = t + kTB * m;
                const uint32_t y = (uint32_t)lane_xor((int)x[m], j, lane);
                const bool take_min = ((i & k) == 0) == ((i & j) == 0);
                x[m] = take_min ? min(x[m], y) : max(x[m], y);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < E; m++) a[t + kTB * m] = x[m];
    __syncthreads();
}

// a[0, n): n a power of two >= 2 and <= 4096; entries [n, kTB) of a are
// overwritten when n < kTB (the buffers hold at least kTB entries)
__device__ void block_sort(uint32_t *a, int n) {
    if (n < kTB) {
        for (int i = n + threadIdx.x; i < kTB; i += kTB) a[i] = 0xffffffffu;
        __syncthreads();
        n = kTB;
    }
    switch (n / kTB) {
        case 1: block_sort_e<1>(a); return;
        case 2: block_sort_e<2>(a); return;
        case 4: block_sort_e<4>(a); return;
        default: break;
    }
    // larger sorts stay in LDS (more entries per thread would cost registers)
    for (int k = 2; k <= n; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < n; i += kTB) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const uint32_t x = a[i], y = a[ixj];
                    if ((x > y) == ((i & k) == 0)) {
                        a[i] = y;
                        a[ixj] = x;
                    }
                }
            }
            __syncthreads();
        }
}

// the number of distinct values of the sorted a[0, n); when it is at most
// cap they are written ascending to out
__device__ int block_unique(const uint32_t *a, int n, int32_t *out, int cap, uint32_t *part) {
    const int t = threadIdx.x;
    const int per = (n + kTB - 1) / kTB;
    const int b = min(n, t * per), e = min(n, b + per);
    uint32_t c = 0;
    for (int i = b; i < e; i++) c += (i == 0 || a[i] != a[i - 1]) ? 1u : 0u;
    part[t] = c;
    scan_parts(part);
    const int total = (int)part[kTB];
    if (total <= cap) {
        uint32_t pos = part[t];
        for (int i = b; i < e; i++)
            if (i == 0 || a[i] != a[i - 1]) out[pos++] = (int32_t)a[i];
    }
    __syncthreads();
    return total;
}

__device__ __forceinline__ int lower_bound_i32(const int32_t *a, int n, int32_t v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Exact sizes of the box [s0, s0 + P) of the Morton order (P <= 255) and
// whether it fits (box_fits).  Leaves in LDS the owned nodes, the sorted
// halo elements and, when H fits, the sorted halo nodes.  All threads of the
// workgroup call it and get the same answer.
__device__ bool eval_box(const MeshView &mv, const int32_t *order, int64_t s0, int P, const TileShape &shape,
                         EvalLds &L, BoxSizes &bs) {
    const int t = threadIdx.x;
    if (t == 0) {
        L.sk = 0;
        L.snt = 0;
    }
    if (t < P) L.owned[t] = order[s0 + t];
    __syncthreads();
    if (t < P) {
        const int32_t v = L.owned[t];
        const int64_t d = mv.n2e_ptr[v + 1] - mv.n2e_ptr[v];
        L.coff[t] = (uint32_t)min<int64_t>(d, kMaxBoxCand + 1);  // a larger degree fails either way
        atomicMax(&L.sk, (int)min<int64_t>(d, 1 << 30));
        atomicAdd(&L.snt, (unsigned long long)(mv.adj_ptr[v + 1] - mv.adj_ptr[v]));
    }
    if (t == P) L.coff[P] = 0;
    __syncthreads();
    const uint32_t cand = block_excl_scan(L.coff, P + 1, L.part);
    bs.P = P;
    bs.cand = cand;
    bs.K = L.sk;
    bs.NT = (int64_t)L.snt;
    bs.NI = (int64_t)cand * mv.npe;
    bs.H = bs.NH = 0;
    if (cand > (uint32_t)kMaxBoxCand || bs.K > 255) return false;
    if (cand == 0) return box_fits(shape, bs);
    if (t < P) {
        const int32_t v = L.owned[t];
        const int64_t q0 = mv.n2e_ptr[v];
        const int d = (int)(mv.n2e_ptr[v + 1] - q0);
        const uint32_t o = L.coff[t];
        for (int j = 0; j < d; j++) L.buf[o + j] = mv.n2e[q0 + j] >> 4;
    }
    int n2 = 2;
    while (n2 < (int)cand) n2 <<= 1;
    for (int i = (int)cand + t; i < n2; i += kTB) L.buf[i] = 0xffffffffu;
    __syncthreads();
    block_sort(L.buf, n2);
    const int hmax = min(shape.max_halo, 255);
    const int H = block_unique(L.buf, (int)cand, L.halo, hmax, L.part);
    bs.H = H;
    if (H > hmax) return false;
    const int nc = H * mv.npe;
    for (int i = t; i < nc; i += kTB) L.buf[i] = (uint32_t)mv.conn[(int64_t)L.halo[i / mv.npe] * mv.npe + i % mv.npe];
    n2 = 2;
    while (n2 < nc) n2 <<= 1;
    for (int i = nc + t; i < n2; i += kTB) L.buf[i] = 0xffffffffu;
    __syncthreads();
    block_sort(L.buf, n2);
    bs.NH = block_unique(L.buf, nc, L.hn, kMaxHaloNodes, L.part);
    return box_fits(shape, bs);
}

// ---- 1. Morton order

__global__ void k_bbox(const double *xyz, int64_t nn, double *part) {
    __shared__ double s[6][256];
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < nn; n += (int64_t)gridDim.x * blockDim.x)
        for (int k = 0; k < 3; k++) {
            const double x = xyz[3 * n + k];
            lo[k] = x < lo[k] ? x : lo[k];
            hi[k] = x > hi[k] ? x : hi[k];
        }
    for (int k = 0; k < 3; k++) {
        s[k][threadIdx.x] = lo[k];
        s[3 + k][threadIdx.x] = hi[k];
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        double r = s[k][0];
        for (int i = 1; i < 256; i++) r = k < 3 ? (s[k][i] < r ? s[k][i] : r) : (s[k][i] > r ? s[k][i] : r);
        part[6 * blockIdx.x + k] = r;
    }
}

// the nb partial boxes of k_bbox reduced to one (one block of 256 threads)
__global__ void k_bbox_final(const double *part, int nb, double *out) {
    __shared__ double s[6][256];
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        for (int k = 0; k < 3; k++) {
            lo[k] = part[6 * b + k] < lo[k] ? part[6 * b + k] : lo[k];
            hi[k] = part[6 * b + 3 + k] > hi[k] ? part[6 * b + 3 + k] : hi[k];
        }
    for (int k = 0; k < 3; k++) {
        s[k][threadIdx.x] = lo[k];
        s[3 + k][threadIdx.x] = hi[k];
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        double r = s[k][0];
        for (int i = 1; i < 256; i++) r = k < 3 ? (s[k][i] < r ? s[k][i] : r) : (s[k][i] > r ? s[k][i] : r);
        out[k] = r;
    }
}

__global__ void k_morton(const double *xyz, int64_t nn, double lo0, double lo1, double lo2, double step, uint64_t *code,
                         int32_t *id) {
    const double lo[3] = {lo0, lo1, lo2};
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < nn; n += (int64_t)gridDim.x * blockDim.x) {
        code[n] = morton_code(xyz + 3 * n, lo, step);
        id[n] = (int32_t)n;
    }
}

// ---- 2. the box tree, one level

// info per tile start: H, NH, K, NT, NI (the sizes box_fits saw)
constexpr int kInfo = 5;

// level l of the tree: the boxes in[0, count[l]) are split or marked as
// tiles; the halves go to out[0, count[l + 1]).  The grid strides over the
// level's boxes, so the host launches levels without knowing their sizes.
__global__ __launch_bounds__(kTB) void k_part_level(const int64_t *in, int64_t *out, unsigned *count, int l,
                                                    unsigned cap, const uint64_t *code, const int32_t *order,
                                                    MeshView mv, TileShape shape, uint8_t *start, int64_t *info,
                                                    unsigned *err) {
    __shared__ EvalLds L;
    const unsigned nin = count[l];
    for (unsigned bx = blockIdx.x; bx < nin; bx += gridDim.x) {
        const int64_t s0 = in[2 * (int64_t)bx], s1 = in[2 * (int64_t)bx + 1];
        const int64_t P = s1 - s0;
        bool fits = false;
        BoxSizes bs;
        if (P <= box_max_nodes(shape)) fits = eval_box(mv, order, s0, (int)P, shape, L, bs);
        if (threadIdx.x == 0) {
            if (fits) {
                start[s0] = 1;
                int64_t *f = info + kInfo * s0;
                f[0] = bs.H;
                f[1] = bs.NH;
                f[2] = bs.K;
                f[3] = bs.NT;
                f[4] = bs.NI;
            } else if (P == 1) {
                atomicOr(err, kErrNodeTooBig);
            } else {
                const int64_t mid = box_split(code, s0, s1);
                const unsigned slot = atomicAdd(&count[l + 1], 2u);
                if (slot + 2 > cap) {
                    atomicOr(err, kErrQueue);
                } else {
                    out[2 * (int64_t)slot] = s0;
                    out[2 * (int64_t)slot + 1] = mid;
                    out[2 * (int64_t)slot + 2] = mid;
                    out[2 * (int64_t)slot + 3] = s1;
                }
            }
        }
        __syncthreads();  // LDS is reused by the next box
    }
}

// tile starts (inc: inclusive scan of the start flags, so that the tile of
// sorted position p is inc[p] - 1)
__global__ void k_tile_starts(const uint8_t *start, const int32_t *inc, int64_t nn, int64_t *ts) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nn; p += (int64_t)gridDim.x * blockDim.x)
        if (start[p]) ts[inc[p] - 1] = p;
    if (blockIdx.x == 0 && threadIdx.x == 0) ts[inc[nn - 1]] = nn;
}

// per tile: halo-node count and slot size (for the scans), maxima for the
// build kernel's LDS
struct PrepStats {
    unsigned long long ntu_max, niu_max, k_max;
};

// largest of v over the workgroup (256 threads), added to *out with one atomic
__device__ void block_max_to(unsigned long long v, unsigned long long *out, unsigned long long *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int d = 128; d > 0; d >>= 1) {
        if ((int)threadIdx.x < d) sh[threadIdx.x] = max(sh[threadIdx.x], sh[threadIdx.x + d]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, sh[0]);
    __syncthreads();
}
__device__ void block_sum_to(unsigned long long v, unsigned long long *out, unsigned long long *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int d = 128; d > 0; d >>= 1) {
        if ((int)threadIdx.x < d) sh[threadIdx.x] += sh[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, sh[0]);
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_tile_prep(const int64_t *ts, int64_t ntiles, const int64_t *info,
                                                   TileShape shape, int64_t *nh, int64_t *slot, PrepStats *st) {
    __shared__ unsigned long long sh[256];
    unsigned long long ntu = 0, niu = 0, km = 0;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t <= ntiles; t += (int64_t)gridDim.x * blockDim.x) {
        if (t == ntiles) {
            nh[t] = 0;
            slot[t] = 0;
            continue;
        }
        const int64_t s0 = ts[t];
        const int P = (int)(ts[t + 1] - s0);
        const int64_t *f = info + kInfo * s0;
        const TileShape::Layout L = shape.layout(f[0], P, f[1], (int)f[2], f[3] + kPadTasks, f[4]);
        nh[t] = f[1];
        slot[t] = (int64_t)(L.U - L.cbase);
        ntu = max(ntu, (unsigned long long)f[3]);
        niu = max(niu, (unsigned long long)f[4]);
        km = max(km, (unsigned long long)f[2]);
    }
    block_max_to(ntu, &st->ntu_max, sh);
    block_max_to(niu, &st->niu_max, sh);
    block_max_to(km, &st->k_max, sh);
}

// ---- 3. one tile

// capacities of the build kernel's dynamic LDS (maxima over the tiles)
struct BuildCaps {
    int ntu;     // unfolded block tasks
    int niu;     // unfolded items
    int nsets;   // bank-spread read sets (< 4096: the read entries are 16-bit)
    int spread;  // H8 FP64 bank-spread numbering
};

constexpr int kSlotsExtra = kPadTasks + 256;  // schedule slots beyond the tasks

// Dynamic LDS of the build kernel.  The task arrays (word .. rows) are dead
// once the records and items are written: the bank spread's load counters
// reuse them, and its read entries reuse the evaluation's sort buffer when
// they fit there.
struct BuildLds {
    size_t word, tpair, sorted, rows, load, bcount, cq, csum, qtab, ibase, lconn, ownc, items, sinfo, gk, occcnt,
        occfill, occ, perm, order, slotl, total;
    bool occ_in_buf;
};

__host__ __device__ inline size_t bl_take(size_t &o, size_t bytes) {
    const size_t r = o;
    o = align16(o + bytes);
    return r;
}

__host__ __device__ inline BuildLds build_lds(const BuildCaps &c, int npe) {
    BuildLds B;
    size_t o = 0;
    const size_t ns = (size_t)c.ntu + kSlotsExtra;  // schedule slots
    const size_t nch = (size_t)c.ntu / 64 + 2;       // chunks of 64 tasks
    const size_t nsg = nch + 8;                      // slot groups
    B.word = bl_take(o, (size_t)c.ntu * 4);
    B.tpair = bl_take(o, (size_t)c.ntu * 2);
    B.sorted = bl_take(o, (size_t)c.ntu * 2);
    B.rows = bl_take(o, (size_t)c.ntu * 2);
    B.load = B.word;
    if (c.spread && (size_t)c.nsets * 16 > o) o = align16((size_t)c.nsets * 16);
    B.bcount = bl_take(o, 513 * 4);
    B.cq = bl_take(o, nch * 4);
    B.csum = bl_take(o, nch * 4);
    B.qtab = bl_take(o, nsg * 4);
    B.ibase = bl_take(o, (nsg + 1) * 4);
    B.lconn = bl_take(o, (size_t)255 * npe * 2);
    B.ownc = bl_take(o, (size_t)255 * npe * 2);
    B.occ_in_buf = (size_t)c.niu * 2 * 2 <= (size_t)kMaxBoxCand * 4;
    if (c.spread) {
        B.items = bl_take(o, (size_t)c.niu * 2);
        B.sinfo = bl_take(o, ns * 4);
        B.gk = bl_take(o, (ns / 16 + 2) * 4);
        B.occcnt = bl_take(o, 257 * 4);
        B.occfill = bl_take(o, 257 * 4);
        B.occ = B.occ_in_buf ? 0 : bl_take(o, (size_t)c.niu * 2 * 2);
        B.perm = bl_take(o, 256 * 4);
        B.order = bl_take(o, 256 * 4);
        B.slotl = bl_take(o, (256 + 32) * 4);
    } else {
        B.items = B.sinfo = B.gk = B.occcnt = B.occfill = B.occ = B.perm = B.order = B.slotl = 0;
    }
    B.total = o;
    return B;
}

// The tile's node incidence lists in LDS: keys[nptr[j] .. nptr[j + 1]) are
// the (halo element h, local index a) of halo node j as h << 4 | a in the low
// 16 bits, ascending (j << 16 in the high bits).  merge_halo walks the
// elements holding both halo nodes ja (column) and jb (row) in ascending
// element, then local index of ja, then of jb -- for the owned column node
// every element is a halo element, so these are all the elements the two
// nodes share -- calling emit(h, ac, am, j) for the j-th; returns their
// number.
template <class F>
__device__ int merge_halo(const uint32_t *keys, const uint16_t *nptr, int ja, int jb, F &&emit) {
    const int a1 = nptr[ja + 1], b1 = nptr[jb + 1];
    int j = nptr[jb], k = 0;
    for (int i = nptr[ja]; i < a1; i++) {
        const uint32_t ce = keys[i] & 0xffffu;
        const uint32_t e = ce >> 4;
        while (j < b1 && ((keys[j] & 0xffffu) >> 4) < e) j++;
        for (int q = j; q < b1 && ((keys[q] & 0xffffu) >> 4) == e; q++) {
            emit(e, ce & 15u, keys[q] & 15u, k);
            k++;
        }
    }
    return k;
}

// the stable-sort key of a kept task: longest first, the symmetric diagonal
// blocks first among equal counts
__device__ __forceinline__ int task_bucket(uint32_t w, int K) {
    return 2 * (K - (int)((w >> 16) & 255u)) + (((w >> 26) & 1u) ? 0 : 1);
}

constexpr uint32_t kDropBit = 1u << 27;  // builder-internal: computed by the transposed block's task

// the owned column whose node block list starts at or before task u
__device__ __forceinline__ int column_of(const uint32_t *coff, int P, int u) {
    int lo = 0, hi = P;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (coff[mid] <= (uint32_t)u)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// the 16 entries of a quarter-wavefront's read row (32-byte aligned), as 8 words
__device__ __forceinline__ void load_row16(const uint16_t *row, uint32_t (&w)[8]) {
    const uint4 a = reinterpret_cast<const uint4 *>(row)[0], b = reinterpret_cast<const uint4 *>(row)[1];
    w[0] = a.x;
    w[1] = a.y;
    w[2] = a.z;
    w[3] = a.w;
    w[4] = b.x;
    w[5] = b.y;
    w[6] = b.z;
    w[7] = b.w;
}
// no entry x < n of the row equals r
__device__ __forceinline__ bool first_in_row(const uint32_t (&w)[8], int n, uint32_t r) {
    bool f = true;
#pragma unroll
    for (int x = 0; x < 16; x++) {
        const uint32_t v = (w[x >> 1] >> ((x & 1) * 16)) & 0xffffu;
        f = f && !(x < n && v == r);
    }
    return f;
}

__global__ __launch_bounds__(kTB) void k_tile_build(MeshView mv, const double *xyz, const int32_t *order,
                                                    const int64_t *ts, const int64_t *info, const int64_t *x0,
                                                    const int64_t *slot_off, TileShape shape, BuildCaps caps,
                                                    unsigned char *slots, int32_t *hnodes, int32_t *tout,
                                                    unsigned *err, unsigned long long *tim) {
    __shared__ __align__(16) EvalLds L;
    __shared__ uint16_t ownj[256];  // halo-node index of each owned column (0xffff: no element)
    __shared__ int sh_fail, sh_K, sh_nch, sh_nsg, sh_NT;
    __shared__ uint32_t sh_NI;
    extern __shared__ __align__(16) unsigned char dyn[];
    const BuildLds B = build_lds(caps, mv.npe);
    uint32_t *word = reinterpret_cast<uint32_t *>(dyn + B.word);
    uint16_t *tpair = reinterpret_cast<uint16_t *>(dyn + B.tpair);
    uint16_t *sorted = reinterpret_cast<uint16_t *>(dyn + B.sorted);
    uint16_t *rows = reinterpret_cast<uint16_t *>(dyn + B.rows);
    uint32_t *bcount = reinterpret_cast<uint32_t *>(dyn + B.bcount);
    int32_t *cq = reinterpret_cast<int32_t *>(dyn + B.cq);
    uint32_t *csum = reinterpret_cast<uint32_t *>(dyn + B.csum);
    int32_t *qtab = reinterpret_cast<int32_t *>(dyn + B.qtab);
    uint32_t *ibase = reinterpret_cast<uint32_t *>(dyn + B.ibase);
    uint16_t *lconn = reinterpret_cast<uint16_t *>(dyn + B.lconn);
    int16_t *ownc = reinterpret_cast<int16_t *>(dyn + B.ownc);  // owned column of each halo node, or -1

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // phase clock (tim != nullptr): thread 0 adds each phase's wall-clock
    // ticks to tim[phase]
    long long t_mark = tim ? wall_clock64() : 0;
    auto mark = [&](int k) {
        if (tim && tid == 0) {
            const long long now = wall_clock64();
            atomicAdd(&tim[k], (unsigned long long)(now - t_mark));
            t_mark = now;
        }
    };
    const int npe = mv.npe;
    const int64_t t = blockIdx.x;
    const int64_t s0 = ts[t];
    const int P = (int)(ts[t + 1] - s0);

    // halo lists (the same evaluation the partition made)
    BoxSizes bs;
    const bool fits = eval_box(mv, order, s0, P, shape, L, bs);
    const int64_t *f = info + kInfo * s0;
    if (!fits || bs.H != f[0] || bs.NH != f[1] || bs.NT != f[3] || bs.NI != f[4] || bs.NT > caps.ntu ||
        bs.NI > caps.niu) {
        if (tid == 0) atomicOr(err, kErrSizes);
        return;
    }
    mark(0);
    const int H = (int)bs.H, NH = (int)bs.NH, NTu = (int)bs.NT;
    const int nc = H * npe;
    for (int i = tid; i < NH; i += kTB) {
        hnodes[x0[t] + i] = L.hn[i];
        ownc[i] = -1;
    }
    for (int i = tid; i < nc; i += kTB) {
        const int32_t v = mv.conn[(int64_t)L.halo[i / npe] * npe + i % npe];
        lconn[i] = (uint16_t)lower_bound_i32(L.hn, NH, v);
    }
    if (tid < P) {
        const int j = lower_bound_i32(L.hn, NH, L.owned[tid]);
        ownj[tid] = (j < NH && L.hn[j] == L.owned[tid]) ? (uint16_t)j : (uint16_t)0xffff;
    }
    if (tid == 0) {
        sh_fail = 0;
        sh_K = 0;
    }
    __syncthreads();
    // node incidence lists (merge_halo)
    uint32_t *keys = L.buf;
    for (int i = tid; i < nc; i += kTB) keys[i] = ((uint32_t)lconn[i] << 16) | ((uint32_t)(i / npe) << 4) | (uint32_t)(i % npe);
    int n2 = 2;
    while (n2 < nc) n2 <<= 1;
    for (int i = nc + tid; i < n2; i += kTB) keys[i] = 0xffffffffu;
    if (tid < P && ownj[tid] != 0xffff) ownc[ownj[tid]] = (int16_t)tid;
    __syncthreads();
    block_sort(keys, n2);
    uint16_t *nptr = reinterpret_cast<uint16_t *>(keys + nc);  // [NH + 1], after the sorted keys
    for (int i = tid; i < nc; i += kTB)
        if (i == 0 || (keys[i] >> 16) != (keys[i - 1] >> 16)) nptr[keys[i] >> 16] = (uint16_t)i;
    if (tid == 0) nptr[NH] = (uint16_t)nc;
    // columns: coff becomes the first task of each owned column
    if (tid < P) {
        const int32_t v = L.owned[tid];
        const int64_t d = mv.adj_ptr[v + 1] - mv.adj_ptr[v];
        if (d > 255) {
            atomicOr(err, kErrDegree);
            sh_fail = 1;
        }
        L.coff[tid] = (uint32_t)d;
    }
    if (tid == P) L.coff[P] = 0;
    __syncthreads();
    if (sh_fail) return;
    mark(1);
    block_excl_scan(L.coff, P + 1, L.part);
    // the row node of every task, as a halo-node index
    for (int u = tid; u < NTu; u += kTB) {
        const int c = column_of(L.coff, P, u), r = u - (int)L.coff[c];
        const int32_t m = mv.adj[mv.adj_ptr[L.owned[c]] + r];
        const int j = lower_bound_i32(L.hn, NH, m);
        if (j >= NH || L.hn[j] != m) {  // a block the tile's elements do not reach
            atomicOr(err, kErrEmptyBlock);
            sh_fail = 1;
        }
        rows[u] = (uint16_t)j;
    }
    __syncthreads();
    if (sh_fail) return;
    mark(2);

    // block tasks in (column, row) order
    for (int u = tid; u < NTu; u += kTB) {
        const int c = column_of(L.coff, P, u), r = u - (int)L.coff[c];
        const int jc = ownj[c], jm = rows[u];
        const bool diag = jm == jc, sw = jm > jc;  // halo-node indices ascend with the node numbers
        const int c2 = ownc[jm];
        const bool same = !diag && c2 >= 0;
        const bool drop = same && c2 < c, pair = same && c2 > c;
        bool sym = diag;
        const int k = drop ? 0 : merge_halo(keys, nptr, jc, jm, [&](uint32_t, uint32_t ac, uint32_t am, int) {
            if (ac != am) sym = false;
        });
        int r2 = 0;
        if (pair) {
            const int b0 = (int)L.coff[c2], dm = (int)(L.coff[c2 + 1] - L.coff[c2]);
            int lo = 0, hi = dm;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (rows[b0 + mid] < jc)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            r2 = lo;
            if (r2 >= dm || rows[b0 + r2] != jc) {
                atomicOr(err, kErrAsym);
                sh_fail = 1;
            }
        }
        if (k > 255) {
            atomicOr(err, kErrShared);
            sh_fail = 1;
        }
        if (!drop && k == 0) {
            atomicOr(err, kErrEmptyBlock);
            sh_fail = 1;
        }
        const int kk = min(k, 255);
        word[u] = (uint32_t)c | ((uint32_t)r << 8) | ((uint32_t)kk << 16) | ((uint32_t)sw << 24) |
                  ((uint32_t)diag << 25) | ((uint32_t)sym << 26) | (drop ? kDropBit : 0u);
        tpair[u] = pair ? (uint16_t)(c2 | (r2 << 8)) : kNoPair;
        if (!drop) atomicMax(&sh_K, kk);
    }
    __syncthreads();
    if (sh_fail) return;
    mark(3);
    const int K = sh_K;

    // stable counting sort of the kept tasks by task_bucket
    const int nb = 2 * K + 2;
    for (int i = tid; i < nb; i += kTB) bcount[i] = 0;
    __syncthreads();
    for (int u = tid; u < NTu; u += kTB) {
        const uint32_t w = word[u];
        if (!(w & kDropBit)) atomicAdd(&bcount[task_bucket(w, K)], 1u);
    }
    __syncthreads();
    const int NTk = (int)block_excl_scan(bcount, nb, L.part);
    // each bucket's tasks in task order, one wavefront per bucket
    for (int b = wave; b < nb; b += kTB / 64) {
        uint32_t pos = bcount[b];
        for (int base = 0; base < NTu; base += 64) {
            const int u = base + lane;
            const uint32_t w = u < NTu ? word[u] : kDropBit;
            const bool hit = !(w & kDropBit) && task_bucket(w, K) == b;
            const uint64_t mask = __ballot(hit);
            if (hit) sorted[pos + __popcll(mask & ((1ull << lane) - 1ull))] = (uint16_t)u;
            pos += (uint32_t)__popcll(mask);
        }
    }
    __syncthreads();
    mark(4);

    // wavefront schedule: chunks of 64 sorted tasks dealt to the 4 SIMDs,
    // the costliest first, each to the least loaded (first on ties); plain
    // order when that needs more than kPadTasks null tasks
    const int nch = (NTk + 63) / 64;
    for (int ch = tid; ch < nch; ch += kTB) {
        uint32_t sum = 0;
        for (int i = ch * 64; i < min(NTk, ch * 64 + 64); i++) sum += (word[sorted[i]] >> 16) & 255u;
        csum[ch] = sum;
    }
    if (tid == 0) {
        double load[4] = {0.0, 0.0, 0.0, 0.0};
        int qn[4] = {0, 0, 0, 0};
        for (int ch = 0; ch < nch; ch++) {
            const double cost = (double)((word[sorted[ch * 64]] >> 16) & 255u) + 2.0;
            int q = 0;
            for (int i = 1; i < 4; i++)
                if (load[i] < load[q]) q = i;
            cq[ch] = (qn[q] << 2) | q;
            qn[q]++;
            load[q] += cost;
        }
        int rounds = 0;
        for (int i = 0; i < 4; i++) rounds = max(rounds, qn[i]);
        if ((int64_t)rounds * 4 * 64 > (int64_t)NTk + kPadTasks) {
            for (int ch = 0; ch < nch; ch++) cq[ch] = ch;
            rounds = (nch + 3) / 4;
        }
        sh_nsg = rounds * 4;
        sh_NT = 0;
    }
    __syncthreads();
    const int nsg = sh_nsg;
    for (int i = tid; i < nsg; i += kTB) qtab[i] = -1;
    __syncthreads();
    for (int ch = tid; ch < nch; ch += kTB) {
        qtab[cq[ch]] = ch;
        atomicMax(&sh_NT, cq[ch] * 64 + min(64, NTk - ch * 64));
    }
    __syncthreads();
    if (tid == 0) {  // item offsets of the slot groups
        uint32_t acc = 0;
        for (int sg = 0; sg < nsg; sg++) {
            ibase[sg] = acc;
            if (qtab[sg] >= 0) acc += csum[qtab[sg]];
        }
        ibase[nsg] = acc;
        sh_NI = acc;
    }
    __syncthreads();
    mark(5);
    const int NT = sh_NT;
    const uint32_t NI = sh_NI;
    if (NI > 65535u) {
        if (tid == 0) atomicOr(err, kErrItems);
        return;
    }
    const TileShape::Layout Lay = shape.layout(H, P, NH, K, NT, NI);
    if (Lay.total > shape.budget) {
        if (tid == 0) atomicOr(err, kErrBudget);
        return;
    }
    unsigned char *blob = slots + slot_off[t];
    uint64_t *rec = reinterpret_cast<uint64_t *>(blob + (Lay.task - Lay.cbase));
    uint16_t *gitem = reinterpret_cast<uint16_t *>(blob + (Lay.item - Lay.cbase));
    uint16_t *items = caps.spread ? reinterpret_cast<uint16_t *>(dyn + B.items) : gitem;
    uint32_t *sinfo = caps.spread ? reinterpret_cast<uint32_t *>(dyn + B.sinfo) : nullptr;

    // task records and items, one slot group per wavefront
    for (int sg = wave; sg < nsg; sg += kTB / 64) {
        const int ch = qtab[sg];
        const int idx = ch * 64 + lane;
        const bool valid = ch >= 0 && idx < NTk;
        const int u = valid ? sorted[idx] : 0;
        const uint32_t w = valid ? word[u] : 0u;
        const uint32_t k = (w >> 16) & 255u;
        uint32_t inc_k = k;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(inc_k, d, 64);
            if (lane >= d) inc_k += y;
        }
        const uint32_t toff = ibase[sg] + inc_k - k;
        const int s = sg * 64 + lane;
        if (s < NT) {
            rec[s] = valid ? ((uint64_t)(w & 0x07ffffffu) | ((uint64_t)toff << 32) | ((uint64_t)tpair[u] << 48))
                           : ((uint64_t)kNullTask | ((uint64_t)kNoPair << 48));
            if (sinfo) sinfo[s] = valid ? (toff | (k << 16)) : 0xffffffffu;
        }
        if (valid) {
            const int c = (int)(w & 255u);
            merge_halo(keys, nptr, ownj[c], rows[u], [&](uint32_t h, uint32_t ac, uint32_t am, int j) {
                items[toff + j] = (uint16_t)((h << 8) | (ac << 4) | am);
            });
        }
    }
    __syncthreads();
    mark(6);

    if (caps.spread && H > 1) {
        // Halo numbering for the LDS banks (the host's spread_halo_banks):
        // read sets per (group of 16 slots, item step j, side), each
        // element's reads, then the greedy class choice, most-read first.
        uint32_t *gk = reinterpret_cast<uint32_t *>(dyn + B.gk);
        uint32_t *occcnt = reinterpret_cast<uint32_t *>(dyn + B.occcnt);
        uint32_t *occfill = reinterpret_cast<uint32_t *>(dyn + B.occfill);
        uint16_t *occ = B.occ_in_buf ? reinterpret_cast<uint16_t *>(L.buf) : reinterpret_cast<uint16_t *>(dyn + B.occ);
        uint8_t *load = reinterpret_cast<uint8_t *>(dyn + B.load);
        int32_t *perm = reinterpret_cast<int32_t *>(dyn + B.perm);
        int32_t *eorder = reinterpret_cast<int32_t *>(dyn + B.order);
        int32_t *slotl = reinterpret_cast<int32_t *>(dyn + B.slotl);  // [16] class starts, then the slots
        const int ngroups = (NT + 15) / 16;
        for (int g = tid; g <= ngroups; g += kTB) {
            uint32_t km = 0;
            for (int s = 16 * g; s < min(NT, 16 * g + 16); s++)
                if (sinfo[s] != 0xffffffffu) km = max(km, sinfo[s] >> 16);
            gk[g] = 2 * km;  // 0 for g == ngroups
        }
        for (int i = tid; i <= H; i += kTB) occcnt[i] = 0;
        __syncthreads();
        const int nsets = (int)block_excl_scan(gk, ngroups + 1, L.part);
        if (nsets > caps.nsets) {
            if (tid == 0) atomicOr(err, kErrSets);
            return;
        }
        for (int i = tid; i < nsets * 16; i += kTB) load[i] = 0;
        // The reads of each set: one quarter-wavefront per group of 16 slots,
        // lane l for slot 16 g + l, one set after the other; a read counts
        // once per set (at its first slot).  Pass 0 counts each element's
        // reads, pass 1 places them.
        uint16_t *rq = reinterpret_cast<uint16_t *>(L.part) + (tid & ~15);  // 16 entries per quarter
        const int q = tid >> 4, ql = tid & 15;
        for (int pass = 0; pass < 2; pass++) {
            if (pass == 1) {
                block_excl_scan(occcnt, H + 1, L.part);
                for (int i = tid; i <= H; i += kTB) occfill[i] = occcnt[i];
                __syncthreads();
            }
            for (int g = q; g < ngroups; g += kTB / 16) {
                const int sl = 16 * g + ql;
                const uint32_t in = sl < NT ? sinfo[sl] : 0xffffffffu;
                const uint32_t kl = in == 0xffffffffu ? 0u : (in >> 16);
                const int g0 = (int)gk[g], ng = (int)(gk[g + 1] - gk[g]);  // the group's 2 kmax sets
                for (int j = 0; j < ng / 2; j++) {  // sets g0 + 2 j (column side), g0 + 2 j + 1 (row side)
                    const bool valid = (uint32_t)j < kl;
                    uint16_t r0 = 0xffff, r1 = 0xffff;
                    if (valid) {
                        const uint16_t it = items[(in & 0xffffu) + j];
                        r0 = (uint16_t)(((it >> 8) << 4) | ((it >> 4) & 15u));
                        r1 = (uint16_t)(((it >> 8) << 4) | (it & 15u));
                    }
                    rq[ql] = r0;
                    rq[256 + ql] = r1;
                    __builtin_amdgcn_wave_barrier();
                    uint32_t w0[8], w1[8];
                    load_row16(rq, w0);
                    load_row16(rq + 256, w1);
                    const bool f0 = valid && first_in_row(w0, ql, r0), f1 = valid && first_in_row(w1, ql, r1);
                    if (pass == 0) {
                        if (f0) atomicAdd(&occcnt[r0 >> 4], 1u);
                        if (f1) atomicAdd(&occcnt[r1 >> 4], 1u);
                    } else {
                        const uint32_t si = (uint32_t)(g0 + 2 * j);
                        if (f0) occ[atomicAdd(&occfill[r0 >> 4], 1u)] = (uint16_t)(si * 16u + (r0 & 15u));
                        if (f1) occ[atomicAdd(&occfill[r1 >> 4], 1u)] = (uint16_t)((si + 1u) * 16u + (r1 & 15u));
                    }
                    __builtin_amdgcn_wave_barrier();
                }
            }
            __syncthreads();
        }
        // elements by read count, descending, stable; slots per residue class
        const int uc = shape.uchunks();
        if (tid < 16) slotl[tid] = 0;
        __syncthreads();
        mark(7);
        for (int e = tid; e < H; e += kTB) {
            const uint32_t ce = occcnt[e + 1] - occcnt[e];
            int rank = 0;
            for (int x = 0; x < H; x++) {
                const uint32_t cx = occcnt[x + 1] - occcnt[x];
                rank += (cx > ce || (cx == ce && x < e)) ? 1 : 0;
            }
            eorder[rank] = e;
            atomicAdd(reinterpret_cast<uint32_t *>(&slotl[(uc * e) & 15]), 1u);  // e doubles as the slot number
        }
        __syncthreads();
        if (tid == 0) {
            int acc = 0;
            for (int c = 0; c < 16; c++) {
                const int n = slotl[c];
                slotl[c] = acc;
                acc += n;
            }
        }
        __syncthreads();
        for (int sl = tid; sl < H; sl += kTB) {
            const int cls = (uc * sl) & 15;
            int rank = 0;
            for (int x = 0; x < sl; x++) rank += ((uc * x) & 15) == cls ? 1 : 0;
            slotl[16 + slotl[cls] + rank] = sl;
        }
        __syncthreads();
        mark(8);
        if (wave == 0) {
            // The four lanes 4 c .. 4 c + 3 hold class c (its first slot, its
            // slot count, the slots used) and sum its cost over every fourth
            // read of the element; the class chosen is the minimum of (cost,
            // -remaining slots, class) over the classes with a slot left.
            const int c = lane >> 2, sub = lane & 3;
            const int cstart = slotl[c];
            const int ccount = (c < 15 ? slotl[c + 1] : H) - cstart;
            int used = 0;
            int e = eorder[0];
            int o0 = (int)occcnt[e], o1 = (int)occcnt[e + 1];
            for (int idx = 0; idx < H; idx++) {
                // the next element's reads (independent of this one's choice)
                const int en = idx + 1 < H ? eorder[idx + 1] : 0;
                const int n0 = (int)occcnt[en], n1 = (int)occcnt[en + 1];
                // this element's reads for the load update below, loaded ahead
                // of the choice (they do not depend on it)
                const uint32_t pre0 = o0 + lane < o1 ? occ[o0 + lane] : 0u;
                const uint32_t pre1 = o0 + lane + 64 < o1 ? occ[o0 + lane + 64] : 0u;
                uint32_t part = 0;
                for (int i0 = o0 + sub; i0 < o1; i0 += 32) {  // eight reads at a time, no branches
                    uint32_t x[8];
#pragma unroll
                    for (int r = 0; r < 8; r++) x[r] = occ[i0 + 4 * r < o1 ? i0 + 4 * r : o0];
#pragma unroll
                    for (int r = 0; r < 8; r++) {
                        const uint32_t v = load[(x[r] >> 4) * 16 + (((x[r] & 15u) + (uint32_t)c) & 15u)];
                        part += i0 + 4 * r < o1 ? v : 0u;
                    }
                }
                // sum over the quad, then the minimum key over the 16 quads
                part += (uint32_t)__builtin_amdgcn_mov_dpp((int)part, 0xB1, 0xf, 0xf, false);  // quad_perm 1,0,3,2
                part += (uint32_t)__builtin_amdgcn_mov_dpp((int)part, 0x4E, 0xf, 0xf, false);  // quad_perm 2,3,0,1
                const int rem = ccount - used;
                uint32_t key = rem > 0 ? (((part & 0xffffu) << 12) | ((uint32_t)(255 - rem) << 4) | (uint32_t)c)
                                       : 0xffffffffu;
                key = min(key, (uint32_t)__builtin_amdgcn_mov_dpp((int)key, 0x141, 0xf, 0xf, false));  // row_half_mirror
                key = min(key, (uint32_t)__builtin_amdgcn_mov_dpp((int)key, 0x140, 0xf, 0xf, false));  // row_mirror
                // every lane of a row of 16 holds the row's minimum: the four rows
                // are combined in scalar registers
                const uint32_t k0 = (uint32_t)__builtin_amdgcn_readlane((int)key, 0);
                const uint32_t k1 = (uint32_t)__builtin_amdgcn_readlane((int)key, 16);
                const uint32_t k2 = (uint32_t)__builtin_amdgcn_readlane((int)key, 32);
                const uint32_t k3 = (uint32_t)__builtin_amdgcn_readlane((int)key, 48);
                const int best = (int)(min(min(k0, k1), min(k2, k3)) & 15u);
                if (c == best) {
                    if (sub == 0) perm[e] = slotl[16 + cstart + used];
                    used++;
                }
                auto bump = [&](uint32_t xx) {
                    uint8_t &ld = load[(xx >> 4) * 16 + (((xx & 15u) + (uint32_t)best) & 15u)];
                    if (ld < 255) ld = (uint8_t)(ld + 1);
                };
                if (o0 + lane < o1) bump(pre0);
                if (o0 + lane + 64 < o1) bump(pre1);
                for (int i = o0 + lane + 128; i < o1; i += 64) bump(occ[i]);
                e = en;
                o0 = n0;
                o1 = n1;
            }
        }
        __syncthreads();
        mark(9);
        for (uint32_t i = tid; i < NI; i += kTB) {
            const uint16_t it = items[i];
            gitem[i] = (uint16_t)((perm[it >> 8] << 8) | (it & 255u));
        }
        uint16_t *glconn = reinterpret_cast<uint16_t *>(blob + (Lay.lconn - Lay.cbase));
        for (int i = tid; i < nc; i += kTB) glconn[perm[i / npe] * npe + i % npe] = lconn[i];
    } else {
        if (caps.spread)
            for (uint32_t i = tid; i < NI; i += kTB) gitem[i] = items[i];
        uint16_t *glconn = reinterpret_cast<uint16_t *>(blob + (Lay.lconn - Lay.cbase));
        for (int i = tid; i < nc; i += kTB) glconn[i] = lconn[i];
    }

    // column pointers, coordinates, the alignment gaps
    int64_t *cbase = reinterpret_cast<int64_t *>(blob);
    int32_t *cdeg = reinterpret_cast<int32_t *>(blob + (Lay.cdeg - Lay.cbase));
    for (int c = tid; c < P; c += kTB) {
        const int32_t v = L.owned[c];
        cbase[c] = mv.adj_ptr[v];
        cdeg[c] = (int32_t)(mv.adj_ptr[v + 1] - mv.adj_ptr[v]);
    }
    double *X = reinterpret_cast<double *>(blob + (Lay.X - Lay.cbase));
    for (int i = tid; i < NH; i += kTB) {
        const int64_t n = L.hn[i];
        X[2 * i] = xyz[3 * n];
        X[2 * i + 1] = xyz[3 * n + 1];
        X[2 * NH + i] = xyz[3 * n + 2];
    }
    if (tid < 6) {
        const size_t ends[6] = {Lay.cbase + (size_t)P * 8,  Lay.cdeg + (size_t)P * 4,
                                Lay.task + (size_t)NT * 8,  Lay.item + (size_t)NI * 2,
                                Lay.lconn + (size_t)nc * 2, Lay.X + (size_t)NH * 24};
        const size_t next[6] = {Lay.cdeg, Lay.task, Lay.item, Lay.lconn, Lay.X, Lay.U};
        for (size_t b = ends[tid]; b < next[tid]; b++) blob[b - Lay.cbase] = 0;
    }
    if (tid == 0) {
        tout[4 * t] = NT;
        tout[4 * t + 1] = K;
        tout[4 * t + 2] = (int32_t)NI;
        tout[4 * t + 3] = (int32_t)(Lay.U - Lay.cbase);
    }
    __syncthreads();
    mark(10);
}

// ---- 4. packing and headers

struct TileStats {
    unsigned long long nhalo, ntasks, nitems, max_nodes, max_halo, max_k;
};

__global__ void k_tile_sizes(const int32_t *tout, int64_t ntiles, int64_t *nbytes) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t <= ntiles; t += (int64_t)gridDim.x * blockDim.x)
        nbytes[t] = t < ntiles ? tout[4 * t + 3] : 0;
}

__global__ void k_tile_pack(const int64_t *ts, const int64_t *info, const int32_t *tout, const int64_t *x0,
                            const int64_t *slot_off, const int64_t *boff, const unsigned char *slots,
                            unsigned char *blob, TileHdr *hdr) {
    const int64_t t = blockIdx.x;
    const int64_t nbytes = tout[4 * t + 3];
    const uint4 *src = reinterpret_cast<const uint4 *>(slots + slot_off[t]);
    uint4 *dst = reinterpret_cast<uint4 *>(blob + boff[t]);
    for (int64_t i = threadIdx.x; i < nbytes / 16; i += blockDim.x) dst[i] = src[i];
    if (threadIdx.x == 0) {
        const int64_t *f = info + kInfo * ts[t];
        TileHdr h;
        h.P = (int32_t)(ts[t + 1] - ts[t]);
        h.H = (int32_t)f[0];
        h.NH = (int32_t)f[1];
        h.NT = tout[4 * t];
        h.K = tout[4 * t + 1];
        h.NI = tout[4 * t + 2];
        h.x0 = (int32_t)x0[t];
        h.nbytes = (int32_t)nbytes;
        h.blob = boff[t];
        hdr[t] = h;
    }
}

__global__ __launch_bounds__(256) void k_tile_stats(const TileHdr *hdr, int64_t ntiles, TileStats *st) {
    __shared__ unsigned long long sh[256];
    unsigned long long nh = 0, nt = 0, ni = 0, mp = 0, mh = 0, mk = 0;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < ntiles; t += (int64_t)gridDim.x * blockDim.x) {
        const TileHdr h = hdr[t];
        nh += h.H;
        nt += h.NT;
        ni += h.NI;
        mp = max(mp, (unsigned long long)h.P);
        mh = max(mh, (unsigned long long)h.H);
        mk = max(mk, (unsigned long long)h.K);
    }
    block_sum_to(nh, &st->nhalo, sh);
    block_sum_to(nt, &st->ntasks, sh);
    block_sum_to(ni, &st->nitems, sh);
    block_max_to(mp, &st->max_nodes, sh);
    block_max_to(mh, &st->max_halo, sh);
    block_max_to(mk, &st->max_k, sh);
}

// (the scan's scratch goes to hold: released by the caller once s is idle)
template <class T>
void scan_excl(hipStream_t s, const T *in, T *out, int64_t n, std::vector<DevPtr> &hold) {
    size_t bytes = 0;
    FEASM_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, (T)0, (size_t)n, rocprim::plus<T>(), s));
    hold.push_back(dev_alloc(std::max<size_t>(bytes, 16)));
    FEASM_HIP(rocprim::exclusive_scan(hold.back()->p, bytes, in, out, (T)0, (size_t)n, rocprim::plus<T>(), s));
}

// FEASM_TILE_TIMING=1 prints the device builder's phase times to stderr
struct Clock {
    bool on = std::getenv("FEASM_TILE_TIMING") != nullptr;
    hipStream_t s;
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    explicit Clock(hipStream_t st) : s(st) {}
    void lap(const char *what) {
        if (!on) return;
        (void)hipStreamSynchronize(s);
        const auto n = std::chrono::steady_clock::now();
        std::fprintf(stderr, "device tiles: %-22s %8.3f ms\n", what,
                     std::chrono::duration<double, std::milli>(n - t).count());
        t = n;
    }
};

void raise_tile_errors(unsigned e) {
    if (!e) return;
    if (e & kErrNodeTooBig)
        throw Error(FEASM_EUNSUPPORTED, "a node has too many incident elements for the fused kernel");
    if (e & kErrDegree) throw Error(FEASM_EUNSUPPORTED, "a node has more than 255 neighbours");
    if (e & kErrShared) throw Error(FEASM_EUNSUPPORTED, "a node block is shared by more than 255 elements");
    if (e & kErrItems) throw Error(FEASM_EUNSUPPORTED, "tile holds too many (block, element) items");
    if (e & kErrAsym) throw Error(FEASM_ESTATE, "pattern is not symmetric");
    if (e & kErrEmptyBlock) throw Error(FEASM_ESTATE, "a pattern block has no element");
    if (e & kErrBudget) throw Error(FEASM_ESTATE, "a tile exceeds the LDS budget it was sized for");
    if (e & kErrSets)
        throw Error(FEASM_EUNSUPPORTED, "a tile's bank-spread read sets exceed the device tile builder's capacity");
    throw Error(FEASM_ESTATE, "device tile builder: inconsistent sizes (code " + std::to_string(e) + ")");
}

}  // namespace

std::shared_ptr<TileSet> build_tiles_device(feasm_mesh *m, const TileShape &shape, hipStream_t stream) {
    feasm_ctx *ctx = m->ctx;
    const hipStream_t s = stream ? stream : ctx->stream;
    const Pattern &pat = *m->pattern;
    const int64_t nn = m->nnodes;
    const int npe = m->npe;
    FEASM_REQUIRE(npe >= 1 && npe <= kMaxNpe, FEASM_EUNSUPPORTED, "element type without a fused kernel");
    FEASM_REQUIRE(nn < INT32_MAX && m->nelem < (int64_t(1) << 28), FEASM_EUNSUPPORTED,
                  "mesh too large for the tile index widths");
    auto T = std::make_shared<TileSet>();
    T->shape = shape;
    if (nn == 0) {
        T->hdr = dev_alloc(16);
        T->blob = dev_alloc(16);
        T->hnodes = dev_alloc(16);
        return T;
    }
    // Temporaries are released only after the stream is idle, also when an
    // error leaves early: `hold` keeps every one of them and is destroyed
    // after `idle`, whose destructor synchronises the stream (the stream
    // need not be one the memory pool fences).
    std::vector<DevPtr> hold;
    struct Idle {
        hipStream_t s;
        ~Idle() { (void)hipStreamSynchronize(s); }
    } idle{s};
    auto tmp = [&](size_t bytes) {
        DevPtr d = dev_alloc(bytes);
        hold.push_back(d);
        return d;
    };
    Clock clk(s);
    clk.lap("stream work before");
    MeshView mv{dp<int32_t>(m->conn), dp<int64_t>(m->n2e_ptr), dp<uint32_t>(m->n2e), dp<int64_t>(pat.adj_ptr),
                dp<int32_t>(pat.adj), npe};
    const double *xyz = dp<double>(m->xyz);

    // 1. Morton order
    const unsigned nbb = grid_for(nn, 256, 1024);
    DevPtr bb = tmp((size_t)(nbb + 1) * 6 * sizeof(double));
    hipLaunchKernelGGL(k_bbox, dim3(nbb), dim3(256), 0, s, xyz, nn, dp<double>(bb));
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(256), 0, s, dp<double>(bb), (int)nbb, dp<double>(bb) + 6 * nbb);
    FEASM_HIP(hipGetLastError());
    double hb[6];
    read_back(s, hb, dp<double>(bb) + 6 * nbb, sizeof(hb));
    const double lo[3] = {hb[0], hb[1], hb[2]}, hi[3] = {hb[3], hb[4], hb[5]};
    const double step = morton_step(lo, hi, nn);
    clk.lap("bounding box");
    DevPtr code_a = tmp(nn * 8), code_b = tmp(nn * 8), id_a = tmp(nn * 4), id_b = tmp(nn * 4);
    hipLaunchKernelGGL(k_morton, dim3(grid_for(nn, 256, 65536)), dim3(256), 0, s, xyz, nn, lo[0], lo[1], lo[2], step,
                       dp<uint64_t>(code_a), dp<int32_t>(id_a));
    FEASM_HIP(hipGetLastError());
    clk.lap("codes");
    const uint64_t *code;
    const int32_t *order;
    {
        rocprim::double_buffer<uint64_t> kb(dp<uint64_t>(code_a), dp<uint64_t>(code_b));
        rocprim::double_buffer<int32_t> vb(dp<int32_t>(id_a), dp<int32_t>(id_b));
        size_t bytes = 0;
        FEASM_HIP(rocprim::radix_sort_pairs(nullptr, bytes, kb, vb, (size_t)nn, 0u, 63u, s));
        DevPtr sort_tmp = tmp(std::max<size_t>(bytes, 16));
        FEASM_HIP(rocprim::radix_sort_pairs(sort_tmp->p, bytes, kb, vb, (size_t)nn, 0u, 63u, s));
        code = kb.current();
        order = vb.current();
    }
    clk.lap("sort");

    // 2. the box tree
    DevPtr start = tmp(nn), info = tmp((size_t)nn * kInfo * 8);
    DevPtr qa = tmp((size_t)(nn + 2) * 16), qb = tmp((size_t)(nn + 2) * 16);
    constexpr int kMaxLevels = 256;  // each level halves boxes at least once: depth < 64 + log2(nn)
    DevPtr ctr = tmp((kMaxLevels + 2) * sizeof(unsigned));  // [error bits, boxes per level...]
    FEASM_HIP(hipMemsetAsync(start->p, 0, nn, s));
    FEASM_HIP(hipMemsetAsync(ctr->p, 0, (kMaxLevels + 2) * sizeof(unsigned), s));
    const int64_t root[2] = {0, nn};
    const unsigned one = 1;
    FEASM_HIP(hipMemcpyAsync(qa->p, root, sizeof(root), hipMemcpyHostToDevice, s));
    unsigned *d_err = dp<unsigned>(ctr), *d_count = dp<unsigned>(ctr) + 1;
    FEASM_HIP(hipMemcpyAsync(d_count, &one, sizeof(one), hipMemcpyHostToDevice, s));
    unsigned hc[2] = {0, 0};
    int levels = 0;
    for (bool more = true; more;) {
        // a batch of levels between read-backs (a level with no boxes costs
        // one launch)
        for (int b = 0; b < 8; b++, levels++) {
            FEASM_REQUIRE(levels < kMaxLevels, FEASM_ESTATE, "device tile builder: box tree too deep");
            const unsigned grid = (unsigned)std::min<int64_t>(std::min<int64_t>(int64_t(1) << std::min(levels, 20), nn), 16384);
            hipLaunchKernelGGL(k_part_level, dim3(grid), dim3(kTB), 0, s, dp<int64_t>(qa), dp<int64_t>(qb), d_count,
                               levels, (unsigned)(nn + 2), code, order, mv, shape, dp<uint8_t>(start),
                               dp<int64_t>(info), d_err);
            FEASM_HIP(hipGetLastError());
            std::swap(qa, qb);
        }
        read_back(s, &hc[0], d_err, sizeof(unsigned));
        read_back(s, &hc[1], d_count + levels, sizeof(unsigned));
        raise_tile_errors(hc[0]);
        more = hc[1] > 0;
    }
    qa.reset();
    qb.reset();
    DevPtr inc = tmp(nn * 4);
    {
        size_t bytes = 0;
        FEASM_HIP(rocprim::inclusive_scan(nullptr, bytes, dp<uint8_t>(start), dp<int32_t>(inc), (size_t)nn,
                                          rocprim::plus<int32_t>(), s));
        DevPtr scan_tmp = tmp(std::max<size_t>(bytes, 16));
        FEASM_HIP(rocprim::inclusive_scan(scan_tmp->p, bytes, dp<uint8_t>(start), dp<int32_t>(inc), (size_t)nn,
                                          rocprim::plus<int32_t>(), s));
    }
    int32_t nt32 = 0;
    read_back(s, &nt32, dp<int32_t>(inc) + nn - 1, 4);
    const int64_t ntiles = nt32;
    DevPtr ts = tmp((ntiles + 1) * 8);
    hipLaunchKernelGGL(k_tile_starts, dim3(grid_for(nn, 256, 65536)), dim3(256), 0, s, dp<uint8_t>(start),
                       dp<int32_t>(inc), nn, dp<int64_t>(ts));
    FEASM_HIP(hipGetLastError());
    clk.lap("partition");

    // sizes, slots, LDS of the build kernel
    DevPtr nh = tmp((ntiles + 1) * 8), slot = tmp((ntiles + 1) * 8);
    DevPtr x0 = tmp((ntiles + 1) * 8), slot_off = tmp((ntiles + 1) * 8);
    DevPtr pst = tmp(sizeof(PrepStats));
    FEASM_HIP(hipMemsetAsync(pst->p, 0, sizeof(PrepStats), s));
    hipLaunchKernelGGL(k_tile_prep, dim3(grid_for(ntiles + 1, 256, 1024)), dim3(256), 0, s, dp<int64_t>(ts), ntiles,
                       dp<int64_t>(info), shape, dp<int64_t>(nh), dp<int64_t>(slot), dp<PrepStats>(pst));
    FEASM_HIP(hipGetLastError());
    scan_excl<int64_t>(s, dp<int64_t>(nh), dp<int64_t>(x0), ntiles + 1, hold);
    scan_excl<int64_t>(s, dp<int64_t>(slot), dp<int64_t>(slot_off), ntiles + 1, hold);
    PrepStats ps;
    int64_t tot[2];
    read_back(s, &ps, pst->p, sizeof(ps));
    read_back(s, &tot[0], dp<int64_t>(x0) + ntiles, 8);
    read_back(s, &tot[1], dp<int64_t>(slot_off) + ntiles, 8);
    FEASM_REQUIRE(tot[0] < INT32_MAX, FEASM_EUNSUPPORTED, "mesh too large for 32-bit tile offsets");
    BuildCaps caps;
    caps.ntu = (int)ps.ntu_max;
    caps.niu = (int)ps.niu_max;
    caps.spread = (shape.grad && shape.rbytes == 8 && npe == 8) ? 1 : 0;
    // sets of the bank spread: per chunk of 64 sorted tasks at most
    // 2 * (K + (chunk items) / 16), the groups' first tasks
    caps.nsets = caps.spread ? std::min(4095, 2 * ((caps.ntu / 64 + 2) * (int)ps.k_max + caps.niu / 16 + 16)) : 0;
    const BuildLds B = build_lds(caps, npe);
    int maxlds = 0;
    FEASM_HIP(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
    FEASM_REQUIRE(B.total + sizeof(EvalLds) + 64 <= (size_t)maxlds, FEASM_EUNSUPPORTED,
                  "tiles too large for the device tile builder");
    T->hnodes = dev_alloc(std::max<int64_t>(tot[0], 1) * 4);
    DevPtr slots = tmp(std::max<int64_t>(tot[1], 16));
    DevPtr tout = tmp(std::max<int64_t>(ntiles, 1) * 16);
    clk.lap("sizes");

    // 3. the tiles
    DevPtr tim;
    if (clk.on) {
        tim = tmp(16 * 8);
        FEASM_HIP(hipMemsetAsync(tim->p, 0, 16 * 8, s));
    }
    if (ntiles > 0) {
        if (B.total > 0)
            FEASM_HIP(hipFuncSetAttribute((const void *)k_tile_build, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)B.total));
        hipLaunchKernelGGL(k_tile_build, dim3((unsigned)ntiles), dim3(kTB), B.total, s, mv, xyz, order,
                           dp<int64_t>(ts), dp<int64_t>(info), dp<int64_t>(x0), dp<int64_t>(slot_off), shape, caps,
                           dp<unsigned char>(slots), dp<int32_t>(T->hnodes), dp<int32_t>(tout), d_err,
                           dp<unsigned long long>(tim));
        FEASM_HIP(hipGetLastError());
    }
    read_back(s, &hc[0], d_err, sizeof(unsigned));
    raise_tile_errors(hc[0]);
    clk.lap("tile arrays");
    if (clk.on) {
        unsigned long long tk[16];
        FEASM_HIP(hipMemcpy(tk, tim->p, sizeof(tk), hipMemcpyDeviceToHost));
        int khz = 0;
        FEASM_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device));
        static const char *names[11] = {"eval",    "lists",  "rows",    "tasks",  "bucket sort", "schedule",
                                        "records", "spread sets", "spread ranks", "spread greedy", "writes"};
        for (int k = 0; k < 11; k++)
            std::fprintf(stderr, "device tiles:   %-16s %10.3f ms summed over tiles\n", names[k],
                         khz ? (double)tk[k] / khz : 0.0);
    }

    // 4. packing
    DevPtr nbytes = tmp((ntiles + 1) * 8), boff = tmp((ntiles + 1) * 8);
    hipLaunchKernelGGL(k_tile_sizes, dim3(grid_for(ntiles + 1, 256, 65536)), dim3(256), 0, s, dp<int32_t>(tout), ntiles,
                       dp<int64_t>(nbytes));
    FEASM_HIP(hipGetLastError());
    scan_excl<int64_t>(s, dp<int64_t>(nbytes), dp<int64_t>(boff), ntiles + 1, hold);
    int64_t nB = 0;
    read_back(s, &nB, dp<int64_t>(boff) + ntiles, 8);
    T->blob = dev_alloc(std::max<int64_t>(nB, 16));
    if (nB < 16) FEASM_HIP(hipMemsetAsync(T->blob->p, 0, 16, s));
    T->hdr = dev_alloc(std::max<size_t>((size_t)ntiles * sizeof(TileHdr), 16));
    DevPtr tst = tmp(sizeof(TileStats));
    FEASM_HIP(hipMemsetAsync(tst->p, 0, sizeof(TileStats), s));
    if (ntiles > 0) {
        hipLaunchKernelGGL(k_tile_pack, dim3((unsigned)ntiles), dim3(256), 0, s, dp<int64_t>(ts), dp<int64_t>(info),
                           dp<int32_t>(tout), dp<int64_t>(x0), dp<int64_t>(slot_off), dp<int64_t>(boff),
                           dp<unsigned char>(slots), dp<unsigned char>(T->blob), dp<TileHdr>(T->hdr));
        FEASM_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_tile_stats, dim3(grid_for(ntiles, 256, 1024)), dim3(256), 0, s, dp<TileHdr>(T->hdr),
                           ntiles, dp<TileStats>(tst));
        FEASM_HIP(hipGetLastError());
    }
    TileStats st;
    read_back(s, &st, tst->p, sizeof(st));
    clk.lap("packing");
    T->ntiles = ntiles;
    T->nhalo = (int64_t)st.nhalo;
    T->ntasks = (int64_t)st.ntasks;
    T->nitems = (int64_t)st.nitems;
    T->blob_bytes = nB;
    T->max_nodes = (int)st.max_nodes;
    T->max_halo = (int)st.max_halo;
    T->max_k = (int)st.max_k;
    if (clk.on)
        std::fprintf(stderr, "device tiles: %lld tiles, %d levels, build LDS %zu + %zu bytes\n", (long long)ntiles,
                     levels, B.total, sizeof(EvalLds));
    return T;
}

}  // namespace feasm
