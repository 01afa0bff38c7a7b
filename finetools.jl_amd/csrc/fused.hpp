// Fused numeric kernel: shared declarations (tile geometry and LDS layout).
#pragma once
#include "common.hpp"

namespace feasm {

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

// What one tile may hold, and the LDS layout the kernel uses for it.  The
// same function sizes the tiles on the host and carves LDS on the device.
struct TileShape {
    int npe = 8, nqp = 8;
    int swp = 9;       // values per (element, node pair) partial: 9 (3x3 M) or 1
    int rbytes = 8;    // sizeof of the element arithmetic type
    bool grad = true;  // forms that need spatial gradients (all but DOT)
    int max_nodes = 64, max_halo = 255, max_cj = 1024;
    int threads = 256;  // workgroup size of the fused kernel (256 or 512)
    size_t budget = 0;  // bytes of LDS available to one workgroup

    struct Layout {
        size_t dN, X, lconn, node, ncj, bp, conn, U, S, cj, P, total;
    };
    // H halo elements, CJ (column, element) pairs, P owned nodes, NH distinct
    // nodes of the halo
    __host__ __device__ Layout layout(int64_t H, int64_t CJ, int P, int64_t NH) const {
        Layout L;
        size_t o = 0;
        L.dN = o;
        o = align16(o + (size_t)nqp * npe * 3 * 8);
        L.X = o;
        o = align16(o + (size_t)NH * 3 * 8);
        L.lconn = o;
        o = align16(o + (size_t)H * npe * 2);
        L.node = o;
        o = align16(o + (size_t)(P + 1) * 4);
        L.ncj = o;
        o = align16(o + (size_t)(P + 1) * 4);
        L.bp = o;
        o = align16(o + (size_t)(P + 1) * 4);
        L.conn = o;
        o = align16(o + (size_t)H * npe * 4);
        L.U = o;
        o = align16(o + (grad ? (size_t)H * nqp * npe * 3 * rbytes : 0));
        L.S = o;
        o = align16(o + (size_t)H * nqp * rbytes);
        L.cj = o;
        o = align16(o + (size_t)CJ * 4);
        L.P = o;
        o = align16(o + (size_t)CJ * npe * swp * rbytes);
        L.total = o;
        return L;
    }
    size_t lds_bytes(int64_t H, int64_t CJ, int P, int64_t NH) const { return layout(H, CJ, P, NH).total; }
    int64_t key() const {
        return (((((int64_t)(npe * 64 + nqp) * 16 + swp) * 16 + rbytes * 2 + (grad ? 1 : 0)) * 256 + max_nodes) *
                    4096 +
                (int64_t)(budget / 1024)) *
                   4 +
               (threads == 512 ? 1 : 0);
    }
};

struct TileSet {
    TileShape shape;
    int64_t ntiles = 0, nhalo = 0;
    int max_nodes = 0, max_halo = 0, max_cj = 0;
    DevPtr node_ptr;  // int32 [ntiles+1]
    DevPtr nodes;     // int32 owned nodes, tile after tile
    DevPtr node_cj;   // int32 per owned node: first cj entry (tile-relative)
    DevPtr halo_ptr;  // int32 [ntiles+1]
    DevPtr halo;      // int32 halo elements, ascending within a tile
    DevPtr cj_ptr;    // int32 [ntiles+1]
    DevPtr cj;        // uint32 (owned-local c << 16) | (halo-local e << 8) | local index of c in e
    DevPtr hnode_ptr; // int32 [ntiles+1]
    DevPtr hnodes;    // int32 distinct nodes of each tile's halo
    DevPtr lconn;     // uint16 per halo element, its nodes as indices into the tile's hnodes
};

std::shared_ptr<TileSet> build_tiles(feasm_mesh *m, const TileShape &shape);

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
    int64_t ntiles = 0;
    size_t budget = 0;
    unsigned grid = 0;
    DevPtr entries, tile_entry;
};
void prepare_batch_launch(feasm_ctx *ctx, BatchLaunch &b, const std::vector<BatchMesh> &items);
void launch_batch(feasm_ctx *ctx, const BatchLaunch &b);
TileShape tile_shape_for(feasm_ctx *ctx, int npe, int nqp, int form, int precision);
void launch_fused(feasm_ctx *ctx, const feasm_mesh *m, const Pattern &p, const TileSet &t, const DevQTable &qt,
                  const FormParams &fp, int precision, double *nzval);

}  // namespace feasm
