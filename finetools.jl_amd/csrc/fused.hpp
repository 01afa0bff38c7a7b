// Fused numeric kernel: shared declarations (tile geometry and LDS layout).
#pragma once
#include "common.hpp"

namespace feasm {

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

// What one tile may hold, and the LDS layout the kernel uses for it.  The
// same function sizes the tiles on the host and carves LDS on the device.
struct TileShape {
    int npe = 8, nqp = 8;
    int rbytes = 8;    // sizeof of the element arithmetic type
    bool grad = true;  // forms that need spatial gradients (all but DOT)
    int max_nodes = 64, max_halo = 255;
    int threads = 256;  // workgroup size of the fused kernel (256, 512 or 1024)
    size_t budget = 0;  // bytes of LDS available to one workgroup

    // A halo element's scaled gradients u (3 values per point and node) sit
    // in LDS as 16-byte chunks, so that the block sums read them with single
    // 16-byte loads: G = 16 / (2 rbytes) points per "xy" chunk (u0, u1 of G
    // points), 2G points per "z" chunk (u2 of 2G points); nqp*npe*3*rbytes/16
    // chunks per record.  Consecutive records are one chunk further apart, so
    // that they start in different LDS banks.  (nqp a multiple of 2G.)
    __host__ __device__ int uchunks() const { return nqp * npe * 3 * rbytes / 16 + 1; }
    // the record stride in elements of the arithmetic type
    __host__ __device__ int ustride() const { return uchunks() * 16 / rbytes; }

    struct Layout {
        size_t tab, X, lconn, cbase, cdeg, task, item, U, sgn, total;
    };
    // doubles of the gradient table per point: (d0, d1) pairs of the npe
    // nodes in one plane [nqp][npe], d2 in a second plane [nqp][npe]
    __host__ __device__ int dstride() const { return npe * 3; }
    // H halo elements, P owned nodes, NH distinct nodes of the halo, K the
    // largest number of elements shared by the two nodes of one block, NT
    // block tasks and NI (block, element) items
    // The tile's metadata and halo coordinates [cbase .. X] are one
    // contiguous blob, stored in HBM exactly as it sits in LDS (offsets
    // relative to cbase).
    __host__ __device__ Layout layout(int64_t H, int P, int64_t NH, int K, int64_t NT, int64_t NI) const {
        Layout L;
        size_t o = 0;
        // N [nqp][npe], dN (two planes, see dstride), the 36 material
        // coefficients (FP64), a flag (in an 8-byte slot), the weights [nqp]
        L.tab = o;
        o = align16(o + ((size_t)nqp * (npe + dstride() + 1) + 37) * 8);
        L.cbase = o;
        o = align16(o + (size_t)P * 8);
        L.cdeg = o;
        o = align16(o + (size_t)P * 4);
        (void)K;
        L.task = o;
        o = align16(o + (size_t)NT * 8);
        L.item = o;
        o = align16(o + (size_t)NI * 2);
        L.lconn = o;
        o = align16(o + (size_t)H * npe * 2);
        L.X = o;  // halo-node coordinates, the blob's last part
        o = align16(o + (size_t)NH * 3 * 8);
        L.U = o;  // end of the blob
        o = align16(o + (grad ? (size_t)H * ustride() * rbytes : 0));
        L.sgn = o;  // grad forms: one byte per (element, point), 1 where J w < 0; DOT: J w
        o = align16(o + (grad ? (size_t)H * nqp : (size_t)H * nqp * rbytes));
        L.total = o;
        return L;
    }
    size_t lds_bytes(int64_t H, int P, int64_t NH, int K, int64_t NT, int64_t NI) const {
        return layout(H, P, NH, K, NT, NI).total;
    }
    int64_t key() const {
        return ((((int64_t)(npe * 64 + nqp) * 16 + rbytes * 2 + (grad ? 1 : 0)) * 256 + max_nodes) * 4096 +
                (int64_t)(budget / 1024)) *
                   2048 +
               threads;
    }
};

// Per tile: its sizes and where its blob and halo node list start.
struct TileHdr {
    int32_t P, H, NH, NT;  // owned columns, halo elements, distinct halo nodes, block tasks
    int32_t K, NI;         // most elements shared by one block's nodes; (block, element) items
    int32_t x0;            // halo nodes: hnodes [x0, x0+NH)
    int32_t nbytes;        // blob size (a multiple of 16), coordinates included
    int64_t blob;          // blob offset in bytes (16-aligned)
};

// Blob contents (see TileShape::layout): int64 cbase[P] (first node block of
// each owned column), int32 cdeg[P] (its node blocks), uint64 task[NT] (one
// record per block task: the task word c | r << 8 | k << 16 | (row > column)
// << 24 | (row == column) << 25 | (a diagonal block none of whose elements
// repeats the node) << 26, or kNullTask; bits 32-47 the task's first item;
// bits 48-63 c2 | r2 << 8 when the task also writes the transposed block, row
// r2 of owned column c2, kNoPair when not), uint16 item[NI] (h << 8 | index
// of the column node in e << 4 | index of the row node; a task's k items in
// ascending element), uint16 lconn[H*npe] (halo connectivity as indices into
// the tile's hnodes), double X[3*NH] (the coordinates of the tile's hnodes:
// NH (x, y) pairs, then NH z values; refresh_tile_coords rewrites them when
// the mesh's coordinates change).
constexpr uint32_t kNullTask = 0xffffffffu;
constexpr uint16_t kNoPair = 0xffff;
// null tasks a tile's wavefront schedule may add (the tile sizing reserves
// room for them)
constexpr int kPadTasks = 320;

struct TileSet {
    TileShape shape;
    int64_t ntiles = 0, nhalo = 0, ntasks = 0, nitems = 0, blob_bytes = 0;
    int max_nodes = 0, max_halo = 0, max_k = 0;
    DevPtr hdr;     // TileHdr [ntiles]
    DevPtr blob;    // the tiles' blobs
    DevPtr hnodes;  // int32 distinct nodes of each tile's halo, ascending
};

// The mesh arrays the tile builder reads (host copies of the device ones):
// 0-based connectivity, coordinates, node -> element lists and the node
// adjacency of the pattern.
struct TileHostMesh {
    int npe = 0;
    int64_t nn = 0, ne = 0;
    std::vector<int32_t> conn;
    std::vector<double> xyz;
    std::vector<int64_t> n2e_ptr, adj_ptr;
    std::vector<uint32_t> n2e;
    std::vector<int32_t> adj;
};
// The tiles as built on the host, before the upload.
struct HostTiles {
    int64_t ntiles = 0, nhalo = 0, ntasks = 0, nitems = 0, blob_bytes = 0;
    int max_nodes = 0, max_halo = 0, max_k = 0;
    std::vector<TileHdr> hdr;
    std::unique_ptr<unsigned char[]> blob;
    size_t blob_size = 0;
    std::vector<int32_t> hnodes;
};
// max_threads bounds the host threads building the per-tile arrays
void build_tiles_host(const TileHostMesh &h, const TileShape &shape, int max_threads, HostTiles &out);
std::shared_ptr<TileSet> build_tiles(feasm_mesh *m, const TileShape &shape, hipStream_t s = nullptr);
// rewrites the coordinates in every tile's blob from the mesh's coordinates
// (on the mesh's context stream)
void refresh_tile_coords(feasm_mesh *m, TileSet &t);

// ---- Tile partition (shared by the host and the device builders, which
// therefore produce the same tiles byte for byte).
//
// Nodes are ordered along a Z-curve (Morton code of their coordinates on a
// lattice whose spacing is the mean node spacing; ties by node number).  A
// box is a run [s0, s1) of that order; the whole order is the first box.  A
// box becomes a tile when its exact sizes fit (box_fits); otherwise it is
// split where the highest code bit that varies inside it flips (both halves
// are then boxes of the Z-curve: box-shaped in space), or in the middle when
// all its codes are equal.  A single node that does not fit is an error.

// 21 bits -> every third bit of a 63-bit Morton code
__host__ __device__ inline uint64_t morton_spread3(uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}
// the lattice spacing for nn nodes in the bounding box [lo, hi] (host only,
// so that both builders quantise with the same number)
double morton_step(const double lo[3], const double hi[3], int64_t nn);
// the Morton code of a point (identical arithmetic on host and device)
__host__ __device__ inline uint64_t morton_code(const double *x, const double lo[3], double step) {
    uint64_t code = 0;
    for (int k = 0; k < 3; k++) {
        const double t = floor((x[k] - lo[k]) / step + 0.5);
        const double c = t < 0.0 ? 0.0 : (t > 2097151.0 ? 2097151.0 : t);
        code |= morton_spread3((uint64_t)c) << k;
    }
    return code;
}

// What decides whether a box fits: owned nodes P, element incidences of the
// owned nodes (cand), distinct halo elements H, distinct halo nodes NH, the
// largest node -> element degree K, and the block tasks NT and items NI
// before transposed blocks are folded (upper bounds of the tile's own).
struct BoxSizes {
    int64_t P = 0, cand = 0, H = 0, NH = 0, K = 0, NT = 0, NI = 0;
};
// most element incidences one box may gather (the device builder sorts
// them in LDS)
constexpr int64_t kMaxBoxCand = 4096;
__host__ __device__ inline int box_max_nodes(const TileShape &s) { return s.max_nodes < 255 ? s.max_nodes : 255; }
__host__ __device__ inline bool box_fits(const TileShape &s, const BoxSizes &b) {
    if (b.P > box_max_nodes(s) || b.cand > kMaxBoxCand || b.K > 255 || b.H > s.max_halo || b.H > 255 ||
        b.NH > 65535)
        return false;
    return s.layout(b.H, (int)b.P, b.NH, (int)b.K, b.NT + kPadTasks, b.NI).total <= s.budget;
}
// where a box that does not fit is cut: the first position whose code has
// the highest varying bit set, or the middle when all codes are equal
// (code: the sorted Morton codes; s1 - s0 >= 2)
__host__ __device__ inline int64_t box_split(const uint64_t *code, int64_t s0, int64_t s1) {
    const uint64_t x = code[s0] ^ code[s1 - 1];
    if (!x) return s0 + (s1 - s0) / 2;
    int b = 63;
    while (!((x >> b) & 1)) b--;
    int64_t lo = s0 + 1, hi = s1 - 1;  // code[s0] has the bit clear, code[s1 - 1] set
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if ((code[mid] >> b) & 1)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// the device builder (csrc/tiles_gpu.hip): the same tiles as
// build_tiles_host, built from the mesh's device arrays on stream s (the
// mesh's context stream when null; the mesh's arrays must be complete)
std::shared_ptr<TileSet> build_tiles_device(feasm_mesh *m, const TileShape &shape, hipStream_t s = nullptr);
// host copies of the arrays the host builder reads
void tile_host_mesh(feasm_mesh *m, TileHostMesh &h);
// builds the tiles with both builders and compares headers, blobs and halo
// node lists byte for byte: the first tile that differs, or -1
int64_t compare_tile_builders(feasm_mesh *m, const TileShape &shape, int64_t *ntiles, double *device_ms);
// the fused kernel's tile shape for this mesh (its tile parameters applied)
TileShape mesh_tile_shape(feasm_mesh *m, int nqp, int form, int precision);

// one mesh of a batched launch
struct BatchMesh {
    feasm_mesh *mesh;
    std::shared_ptr<TileSet> tiles;
    FormParams fp;
    int rule, order;
    double *nzval;
};
// a prepared persistent launch over the tiles of several meshes that share
// the element type, rule size and arithmetic
struct BatchLaunch {
    int npe = 8, nqp = 8, rbytes = 8, threads = 256;
    int kform = 0, ndn = 1;  // kernel form (fused_kernel_form) and DOFs per node
    int64_t ntiles = 0;
    size_t budget = 0;
    unsigned grid = 0;
    DevPtr entries, tile_entry, hdr;  // hdr: the tiles' headers, mesh after mesh
    TileShape shape;  // the meshes' tile layout (npe, nqp, rbytes and grad agree)
};
void prepare_batch_launch(feasm_ctx *ctx, BatchLaunch &b, const std::vector<BatchMesh> &items);
// the fused kernel's form id for these parameters (forms with a specialised
// kernel map to their own id)
int fused_kernel_form(const FormParams &fp);
void launch_batch(feasm_ctx *ctx, const BatchLaunch &b);
TileShape tile_shape_for(feasm_ctx *ctx, int npe, int nqp, int form, int precision);
void launch_fused(feasm_ctx *ctx, const feasm_mesh *m, const Pattern &p, const TileSet &t, const DevQTable &qt,
                  const FormParams &fp, int precision, double *nzval);

}  // namespace feasm
