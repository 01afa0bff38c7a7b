// Tiles for the fused numeric kernel (symbolic phase, host side).
//
// A tile is a set of "owned" column nodes, contiguous in a Morton (Z-curve)
// order of the node coordinates so that it is spatially compact, together
// with its halo: every element incident to an owned node, in ascending
// element number.  One workgroup integrates the halo elements into LDS and
// produces every CSC block of the owned columns, summing the element
// contributions of each block in ascending element order.  Tiles are grown
// greedily along the Morton order until the next node would overflow the
// LDS budget the kernel has for this form and precision.
#include <algorithm>
#include <numeric>

#include "common.hpp"
#include "fused.hpp"

namespace feasm {

namespace {

uint64_t spread3(uint64_t v) {  // 21 bits -> every third bit
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

}  // namespace

std::shared_ptr<TileSet> build_tiles(feasm_mesh *m, const TileShape &shape) {
    feasm_ctx *ctx = m->ctx;
    const int64_t nn = m->nnodes, ne = m->nelem;
    const int npe = m->npe;
    std::vector<int32_t> conn(ne * npe);
    std::vector<double> xyz(nn * 3);
    std::vector<int64_t> n2e_ptr(nn + 1);
    std::vector<uint32_t> n2e(ne * npe);
    FEASM_HIP(hipMemcpy(conn.data(), m->conn->p, conn.size() * 4, hipMemcpyDeviceToHost));
    FEASM_HIP(hipMemcpy(xyz.data(), m->xyz->p, xyz.size() * 8, hipMemcpyDeviceToHost));
    FEASM_HIP(hipMemcpy(n2e_ptr.data(), m->n2e_ptr->p, (nn + 1) * 8, hipMemcpyDeviceToHost));
    FEASM_HIP(hipMemcpy(n2e.data(), m->n2e->p, n2e.size() * 4, hipMemcpyDeviceToHost));

    // Morton order of the nodes
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t n = 0; n < nn; n++)
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], xyz[3 * n + k]);
            hi[k] = std::max(hi[k], xyz[3 * n + k]);
        }
    double span = 0;
    for (int k = 0; k < 3; k++) span = std::max(span, hi[k] - lo[k]);
    if (!(span > 0)) span = 1;
    std::vector<std::pair<uint64_t, int32_t>> key(nn);
    for (int64_t n = 0; n < nn; n++) {
        uint64_t code = 0;
        for (int k = 0; k < 3; k++) {
            double t = (xyz[3 * n + k] - lo[k]) / span;
            uint64_t q = (uint64_t)std::min(2097151.0, std::max(0.0, t * 2097151.0));
            code |= spread3(q) << k;
        }
        key[n] = {code, (int32_t)n};
    }
    std::sort(key.begin(), key.end());

    auto T = std::make_shared<TileSet>();
    T->shape = shape;
    std::vector<int32_t> node_ptr{0}, nodes, halo_ptr{0}, halo, cj_ptr{0}, node_cj, hnode_ptr{0}, hnodes;
    std::vector<uint32_t> cj;
    std::vector<uint16_t> lconn;
    std::vector<int32_t> nstamp(nn, -1), nlocal(nn, 0);  // halo-node membership of the current tile
    int64_t cur_nh = 0;
    nodes.reserve(nn);
    node_cj.reserve(nn + 1);
    std::vector<int32_t> stamp(ne, -1), local(ne, 0);
    std::vector<int32_t> cur_nodes, cur_halo;
    int64_t cur_cj = 0;
    int32_t tile = 0;
    auto close_tile = [&]() {
        if (cur_nodes.empty()) return;
        std::sort(cur_halo.begin(), cur_halo.end());
        for (size_t i = 0; i < cur_halo.size(); i++) local[cur_halo[i]] = (int32_t)i;
        int32_t cstart = 0;
        for (size_t c = 0; c < cur_nodes.size(); c++) {
            const int32_t n = cur_nodes[c];
            nodes.push_back(n);
            node_cj.push_back(cstart);
            for (int64_t s = n2e_ptr[n]; s < n2e_ptr[n + 1]; s++) {  // ascending element
                const uint32_t e = n2e[s] >> 4, a = n2e[s] & 15u;
                cj.push_back(((uint32_t)c << 16) | ((uint32_t)local[e] << 8) | a);
                cstart++;
            }
        }
        halo.insert(halo.end(), cur_halo.begin(), cur_halo.end());
        // distinct halo nodes, ascending, and the halo connectivity in local indices
        std::vector<int32_t> hn;
        for (int32_t e : cur_halo)
            for (int a = 0; a < npe; a++) hn.push_back(conn[(int64_t)e * npe + a]);
        std::sort(hn.begin(), hn.end());
        hn.erase(std::unique(hn.begin(), hn.end()), hn.end());
        for (size_t i = 0; i < hn.size(); i++) nlocal[hn[i]] = (int32_t)i;
        for (int32_t e : cur_halo)
            for (int a = 0; a < npe; a++) lconn.push_back((uint16_t)nlocal[conn[(int64_t)e * npe + a]]);
        hnodes.insert(hnodes.end(), hn.begin(), hn.end());
        hnode_ptr.push_back((int32_t)hnodes.size());
        node_ptr.push_back((int32_t)nodes.size());
        halo_ptr.push_back((int32_t)halo.size());
        cj_ptr.push_back((int32_t)cj.size());
        T->max_nodes = std::max<int>(T->max_nodes, (int)cur_nodes.size());
        T->max_halo = std::max<int>(T->max_halo, (int)cur_halo.size());
        T->max_cj = std::max<int>(T->max_cj, (int)cur_cj);
        cur_nodes.clear();
        cur_halo.clear();
        cur_cj = 0;
        cur_nh = 0;
        tile++;
    };
    std::vector<int32_t> seen;
    seen.reserve(256);
    for (int64_t i = 0; i < nn; i++) {
        const int32_t n = key[i].second;
        const int64_t deg = n2e_ptr[n + 1] - n2e_ptr[n];
        // what adding node n would add: new halo elements and new halo nodes
        auto count_new = [&](int64_t &fresh, int64_t &fresh_nodes) {
            fresh = 0;
            fresh_nodes = 0;
            seen.clear();
            for (int64_t s = n2e_ptr[n]; s < n2e_ptr[n + 1]; s++) {
                const int32_t e = (int32_t)(n2e[s] >> 4);
                if (stamp[e] == tile) continue;
                fresh++;
                for (int a = 0; a < npe; a++) {
                    const int32_t v = conn[(int64_t)e * npe + a];
                    if (nstamp[v] != tile && std::find(seen.begin(), seen.end(), v) == seen.end()) {
                        seen.push_back(v);
                        fresh_nodes++;
                    }
                }
            }
        };
        int64_t fresh, fresh_nodes;
        count_new(fresh, fresh_nodes);
        const int64_t H = (int64_t)cur_halo.size() + fresh, CJ = cur_cj + deg, NH = cur_nh + fresh_nodes;
        const bool fits = (int64_t)cur_nodes.size() + 1 <= shape.max_nodes && H <= shape.max_halo &&
                          CJ <= shape.max_cj && NH <= 65535 &&
                          shape.lds_bytes(H, CJ, (int)cur_nodes.size() + 1, NH) <= shape.budget;
        if (!fits && !cur_nodes.empty()) {
            close_tile();
            count_new(fresh, fresh_nodes);
        }
        FEASM_REQUIRE(deg <= shape.max_cj && deg <= shape.max_halo &&
                          shape.lds_bytes(deg, deg, 1, fresh_nodes) <= shape.budget,
                      FEASM_EUNSUPPORTED, "a node has too many incident elements for the fused kernel");
        for (int64_t s = n2e_ptr[n]; s < n2e_ptr[n + 1]; s++) {
            const int32_t e = (int32_t)(n2e[s] >> 4);
            if (stamp[e] != tile) {
                stamp[e] = tile;
                cur_halo.push_back(e);
                for (int a = 0; a < npe; a++) {
                    const int32_t v = conn[(int64_t)e * npe + a];
                    if (nstamp[v] != tile) {
                        nstamp[v] = tile;
                        cur_nh++;
                    }
                }
            }
        }
        cur_nodes.push_back(n);
        cur_cj += deg;
    }
    close_tile();
    T->ntiles = tile;
    T->nhalo = (int64_t)halo.size();
    node_cj.push_back(0);  // keeps the array non-empty
    auto up = [&](const void *src, size_t bytes) {
        DevPtr d = dev_alloc(std::max<size_t>(bytes, 4));
        if (bytes) FEASM_HIP(hipMemcpy(d->p, src, bytes, hipMemcpyHostToDevice));
        return d;
    };
    T->node_ptr = up(node_ptr.data(), node_ptr.size() * 4);
    T->nodes = up(nodes.data(), nodes.size() * 4);
    T->node_cj = up(node_cj.data(), node_cj.size() * 4);
    T->halo_ptr = up(halo_ptr.data(), halo_ptr.size() * 4);
    T->halo = up(halo.data(), halo.size() * 4);
    T->cj_ptr = up(cj_ptr.data(), cj_ptr.size() * 4);
    T->cj = up(cj.data(), cj.size() * 4);
    T->hnode_ptr = up(hnode_ptr.data(), hnode_ptr.size() * 4);
    T->hnodes = up(hnodes.data(), hnodes.size() * 4);
    T->lconn = up(lconn.data(), lconn.size() * 2);
    (void)ctx;
    return T;
}

}  // namespace feasm
