// CPU check of the host tile builder (finetools.jl_amd/csrc/tiles.cpp), the
// reference the device builder is compared with byte for byte on the GPU.
//
// For H8 and T10 blocks (perturbed), a slab whose nodes share Morton codes,
// and the FP64 / mixed / DOT tile shapes, it
//  - recomputes the Morton box-tree partition independently (std::set for
//    the halo sizes) and compares the tile boundaries;
//  - checks every tile against the mesh: owned columns and their pattern
//    rows, the halo elements (found back from the halo connectivity) and the
//    halo nodes, every block task's items (the elements holding both nodes,
//    ascending, with their local indices), the symmetric-diagonal and
//    transposed-pair fields, that every node block of every owned column is
//    produced exactly once, the coordinates, the padding bound and the LDS
//    layout;
//  - checks that the bytes do not depend on the number of host threads.
// Prints "ok" and exits 0 when every check passes.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <functional>
#include <set>
#include <vector>

#include "fused.hpp"

namespace feasm {
void gen_h8block(double L, double W, double H, int64_t nL, int64_t nW, int64_t nH, int64_t *conn, double *xyz);
int64_t gen_t10block(double L, double W, double H, int64_t nL, int64_t nW, int64_t nH, double amp_v, double amp_m,
                     uint64_t seed, int64_t *conn10, std::vector<double> &xyz10);
void perturb_nodes(int64_t nnodes, double *xyz, const double lo[3], const double hi[3], double amp, uint64_t seed);
// this program builds on the host only
std::shared_ptr<TileSet> build_tiles_device(feasm_mesh *, const TileShape &, hipStream_t) {
    throw Error(FEASM_EUNSUPPORTED, "no device in the tile check");
}
}  // namespace feasm

using namespace feasm;

static int failures = 0;
#define CHECK(cond, ...)                                   \
    do {                                                   \
        if (!(cond)) {                                     \
            std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__);             \
            std::fprintf(stderr, "\n");                    \
            if (++failures > 20) std::exit(1);             \
            return;                                        \
        }                                                  \
    } while (0)

static void finish_mesh(TileHostMesh &h, const std::vector<int64_t> &conn1) {
    const int npe = h.npe;
    h.conn.resize(h.ne * npe);
    for (int64_t i = 0; i < h.ne * npe; i++) h.conn[i] = (int32_t)(conn1[i] - 1);
    h.n2e_ptr.assign(h.nn + 1, 0);
    for (int64_t i = 0; i < h.ne * npe; i++) h.n2e_ptr[h.conn[i] + 1]++;
    for (int64_t i = 0; i < h.nn; i++) h.n2e_ptr[i + 1] += h.n2e_ptr[i];
    h.n2e.resize(h.ne * npe);
    std::vector<int64_t> cur(h.n2e_ptr.begin(), h.n2e_ptr.end() - 1);
    for (int64_t e = 0; e < h.ne; e++)
        for (int a = 0; a < npe; a++) h.n2e[cur[h.conn[e * npe + a]]++] = (uint32_t)((e << 4) | a);
    h.adj_ptr.assign(h.nn + 1, 0);
    h.adj.clear();
    for (int64_t v = 0; v < h.nn; v++) {
        std::set<int32_t> l;
        for (int64_t s = h.n2e_ptr[v]; s < h.n2e_ptr[v + 1]; s++)
            for (int a = 0; a < npe; a++) l.insert(h.conn[(h.n2e[s] >> 4) * npe + a]);
        h.adj.insert(h.adj.end(), l.begin(), l.end());
        h.adj_ptr[v + 1] = (int64_t)h.adj.size();
    }
}

static TileHostMesh h8(int nx, int ny, int nz, double thick, double amp, uint64_t seed) {
    TileHostMesh h;
    h.npe = 8;
    h.ne = (int64_t)nx * ny * nz;
    h.nn = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    std::vector<int64_t> c(h.ne * 8);
    h.xyz.resize(h.nn * 3);
    gen_h8block(1.0, 1.0, thick, nx, ny, nz, c.data(), h.xyz.data());
    if (amp > 0) {
        const double lo[3] = {0, 0, 0}, hi[3] = {1, 1, thick};
        perturb_nodes(h.nn, h.xyz.data(), lo, hi, amp / std::max(nx, std::max(ny, nz)), seed);
    }
    finish_mesh(h, c);
    return h;
}

// an H8 block of which only the elements with centroid x < 0.45 are kept,
// over all the block's nodes: the others are nodes of no element (a region
// of a larger mesh assembled on the global node set)
static TileHostMesh h8_region(int nx, int ny, int nz) {
    TileHostMesh h;
    h.npe = 8;
    const int64_t ne = (int64_t)nx * ny * nz;
    h.nn = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    std::vector<int64_t> c(ne * 8), kept;
    h.xyz.resize(h.nn * 3);
    gen_h8block(1.0, 1.0, 1.0, nx, ny, nz, c.data(), h.xyz.data());
    for (int64_t e = 0; e < ne; e++) {
        double cx = 0;
        for (int a = 0; a < 8; a++) cx += h.xyz[3 * (c[e * 8 + a] - 1)] / 8;
        if (cx < 0.45) kept.insert(kept.end(), c.begin() + e * 8, c.begin() + e * 8 + 8);
    }
    h.ne = (int64_t)kept.size() / 8;
    finish_mesh(h, kept);
    return h;
}

static TileHostMesh t10(int n, uint64_t seed) {
    TileHostMesh h;
    h.npe = 10;
    h.ne = 6LL * n * n * n;
    std::vector<int64_t> c(h.ne * 10);
    h.nn = gen_t10block(1.0, 1.0, 1.0, n, n, n, 0.1 / n, 0.03 / n, seed, c.data(), h.xyz);
    finish_mesh(h, c);
    return h;
}

// the Morton order, restated: codes from the shared morton_code, sorted
// with the node number
static std::vector<std::pair<uint64_t, int32_t>> order_of(const TileHostMesh &h) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t v = 0; v < h.nn; v++)
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], h.xyz[3 * v + k]);
            hi[k] = std::max(hi[k], h.xyz[3 * v + k]);
        }
    const double step = morton_step(lo, hi, h.nn);
    std::vector<std::pair<uint64_t, int32_t>> key(h.nn);
    for (int64_t v = 0; v < h.nn; v++) key[v] = {morton_code(&h.xyz[3 * v], lo, step), (int32_t)v};
    std::sort(key.begin(), key.end());
    return key;
}

// the box-tree partition, restated with sets: tile start positions
static std::vector<int64_t> partition_of(const TileHostMesh &h, const TileShape &shape,
                                         const std::vector<std::pair<uint64_t, int32_t>> &key) {
    std::vector<uint64_t> code(key.size());
    for (size_t i = 0; i < key.size(); i++) code[i] = key[i].first;
    std::vector<int64_t> starts;
    std::function<void(int64_t, int64_t)> rec = [&](int64_t s0, int64_t s1) {
        BoxSizes b;
        b.P = s1 - s0;
        if (b.P <= box_max_nodes(shape)) {
            std::set<int32_t> E, N;
            for (int64_t i = s0; i < s1; i++) {
                const int32_t v = key[i].second;
                const int64_t d = h.n2e_ptr[v + 1] - h.n2e_ptr[v];
                b.cand += d;
                b.K = std::max(b.K, d);
                b.NT += h.adj_ptr[v + 1] - h.adj_ptr[v];
                for (int64_t q = h.n2e_ptr[v]; q < h.n2e_ptr[v + 1]; q++) E.insert((int32_t)(h.n2e[q] >> 4));
            }
            for (int32_t e : E)
                for (int a = 0; a < h.npe; a++) N.insert(h.conn[(int64_t)e * h.npe + a]);
            b.H = (int64_t)E.size();
            b.NH = (int64_t)N.size();
            b.NI = b.cand * h.npe;
        }
        if (box_fits(shape, b)) {
            starts.push_back(s0);
            return;
        }
        if (s1 - s0 < 2) {
            starts.push_back(-1);  // a node that cannot be a tile
            return;
        }
        const int64_t mid = box_split(code.data(), s0, s1);
        rec(s0, mid);
        rec(mid, s1);
    };
    if (h.nn > 0) rec(0, h.nn);
    return starts;
}

static void check_mesh(const char *name, const TileHostMesh &h, const TileShape &shape) {
    HostTiles T;
    build_tiles_host(h, shape, 8, T);
    const auto key = order_of(h);
    const std::vector<int64_t> starts = partition_of(h, shape, key);
    CHECK((int64_t)starts.size() == T.ntiles, "%s: %lld tiles, the restated partition has %zu", name,
          (long long)T.ntiles, starts.size());
    const int npe = h.npe;
    int64_t s0 = 0, nhalo = 0, ntasks = 0, nitems = 0;
    for (int64_t t = 0; t < T.ntiles; t++) {
        const TileHdr &d = T.hdr[t];
        CHECK(starts[t] == s0, "%s tile %lld: starts at %lld, restated %lld", name, (long long)t, (long long)s0,
              (long long)starts[t]);
        const TileShape::Layout L = shape.layout(d.H, d.P, d.NH, d.K, d.NT, d.NI);
        CHECK(L.total <= shape.budget && d.nbytes == (int32_t)(L.U - L.cbase), "%s tile %lld: layout", name,
              (long long)t);
        const unsigned char *b = T.blob.get() + d.blob;
        const int64_t *cbase = reinterpret_cast<const int64_t *>(b);
        const int32_t *cdeg = reinterpret_cast<const int32_t *>(b + (L.cdeg - L.cbase));
        const uint64_t *task = reinterpret_cast<const uint64_t *>(b + (L.task - L.cbase));
        const uint16_t *item = reinterpret_cast<const uint16_t *>(b + (L.item - L.cbase));
        const uint16_t *lconn = reinterpret_cast<const uint16_t *>(b + (L.lconn - L.cbase));
        const double *X = reinterpret_cast<const double *>(b + (L.X - L.cbase));
        const int32_t *hn = T.hnodes.data() + d.x0;
        std::vector<int32_t> owned(d.P);
        for (int c = 0; c < d.P; c++) {
            owned[c] = key[s0 + c].second;
            CHECK(cbase[c] == h.adj_ptr[owned[c]] && cdeg[c] == h.adj_ptr[owned[c] + 1] - h.adj_ptr[owned[c]],
                  "%s tile %lld: column %d", name, (long long)t, c);
        }
        // halo nodes and halo elements (found back from the connectivity)
        std::set<int32_t> E, N;
        for (int32_t v : owned)
            for (int64_t q = h.n2e_ptr[v]; q < h.n2e_ptr[v + 1]; q++) E.insert((int32_t)(h.n2e[q] >> 4));
        for (int32_t e : E)
            for (int a = 0; a < npe; a++) N.insert(h.conn[(int64_t)e * npe + a]);
        CHECK((int64_t)E.size() == d.H && (int64_t)N.size() == d.NH, "%s tile %lld: halo sizes", name, (long long)t);
        CHECK(std::equal(N.begin(), N.end(), hn), "%s tile %lld: halo nodes", name, (long long)t);
        std::vector<int32_t> elem(d.H);
        std::set<int32_t> seen;
        for (int s = 0; s < d.H; s++) {
            const int32_t v0 = hn[lconn[s * npe]];
            int32_t found = -1;
            for (int64_t q = h.n2e_ptr[v0]; q < h.n2e_ptr[v0 + 1] && found < 0; q++) {
                const int32_t e = (int32_t)(h.n2e[q] >> 4);
                bool same = true;
                for (int a = 0; a < npe; a++) same = same && h.conn[(int64_t)e * npe + a] == hn[lconn[s * npe + a]];
                if (same) found = e;
            }
            CHECK(found >= 0 && E.count(found) && !seen.count(found), "%s tile %lld: halo slot %d", name,
                  (long long)t, s);
            seen.insert(found);
            elem[s] = found;
        }
        for (int i = 0; i < d.NH; i++)
            CHECK(X[2 * i] == h.xyz[3 * hn[i]] && X[2 * i + 1] == h.xyz[3 * hn[i] + 1] &&
                      X[2 * d.NH + i] == h.xyz[3 * hn[i] + 2],
                  "%s tile %lld: coordinates", name, (long long)t);
        // tasks: items, flags, coverage of every node block
        std::vector<std::vector<int>> covered(d.P);
        for (int c = 0; c < d.P; c++) covered[c].assign(cdeg[c], 0);
        int kept = 0, kmax = 0;
        int64_t ni = 0;
        for (int s = 0; s < d.NT; s++) {
            const uint64_t rec = task[s];
            if ((uint32_t)rec == kNullTask) {
                CHECK((rec >> 32) == ((uint64_t)kNoPair << 16), "%s tile %lld: null task %d", name, (long long)t, s);
                continue;
            }
            kept++;
            const uint32_t w = (uint32_t)rec;
            const int c = (int)(w & 255u), r = (int)((w >> 8) & 255u), k = (int)((w >> 16) & 255u);
            const bool sw = (w >> 24) & 1u, diag = (w >> 25) & 1u, sym = (w >> 26) & 1u;
            const int toff = (int)((rec >> 32) & 0xffffu), pair = (int)(rec >> 48);
            CHECK(c < d.P && r < cdeg[c] && (w >> 27) == 0, "%s tile %lld: task %d fields", name, (long long)t, s);
            const int32_t v = owned[c], m = h.adj[cbase[c] + r];
            CHECK(sw == (m > v) && diag == (m == v), "%s tile %lld: task %d orientation", name, (long long)t, s);
            // the elements holding both nodes, ascending, with local indices
            std::vector<std::tuple<int32_t, int, int>> want;
            for (int64_t q = h.n2e_ptr[v]; q < h.n2e_ptr[v + 1]; q++) {
                const int32_t e = (int32_t)(h.n2e[q] >> 4);
                const int ac = (int)(h.n2e[q] & 15u);
                for (int a = 0; a < npe; a++)
                    if (h.conn[(int64_t)e * npe + a] == m) want.emplace_back(e, ac, a);
            }
            CHECK((int)want.size() == k && k > 0, "%s tile %lld: task %d has %d items, %zu elements", name,
                  (long long)t, s, k, want.size());
            bool allsame = true;
            for (int j = 0; j < k; j++) {
                const uint16_t it = item[toff + j];
                const int hl = it >> 8, ac = (it >> 4) & 15, am = it & 15;
                CHECK(hl < d.H && elem[hl] == std::get<0>(want[j]) && ac == std::get<1>(want[j]) &&
                          am == std::get<2>(want[j]),
                      "%s tile %lld: task %d item %d", name, (long long)t, s, j);
                allsame = allsame && ac == am;
            }
            CHECK(sym == (diag && allsame), "%s tile %lld: task %d symmetric flag", name, (long long)t, s);
            covered[c][r]++;
            if (pair != kNoPair) {
                const int c2 = pair & 255, r2 = pair >> 8;
                CHECK(c2 < d.P && c2 != c && owned[c2] == m && r2 < cdeg[c2] && h.adj[cbase[c2] + r2] == v,
                      "%s tile %lld: task %d transposed block", name, (long long)t, s);
                covered[c2][r2]++;
            } else {
                // an unpaired block's row node is not another owned column
                for (int c2 = 0; c2 < d.P; c2++)
                    CHECK(c2 == c || owned[c2] != m, "%s tile %lld: task %d not folded", name, (long long)t, s);
            }
            ni += k;
            kmax = std::max(kmax, k);
        }
        for (int c = 0; c < d.P; c++)
            for (int r = 0; r < cdeg[c]; r++)
                CHECK(covered[c][r] == 1, "%s tile %lld: block (%d, %d) produced %d times", name, (long long)t, c, r,
                      covered[c][r]);
        CHECK(ni == d.NI && kmax == d.K && d.NT <= kept + kPadTasks, "%s tile %lld: counts", name, (long long)t);
        nhalo += d.H;
        ntasks += d.NT;
        nitems += d.NI;
        s0 += d.P;
    }
    CHECK(s0 == h.nn && nhalo == T.nhalo && ntasks == T.ntasks && nitems == T.nitems, "%s: totals", name);
    // the same bytes with one host thread
    HostTiles T1;
    build_tiles_host(h, shape, 1, T1);
    CHECK(T1.ntiles == T.ntiles && T1.blob_bytes == T.blob_bytes &&
              std::memcmp(T1.blob.get(), T.blob.get(), (size_t)T.blob_bytes) == 0 &&
              std::memcmp(T1.hdr.data(), T.hdr.data(), T.hdr.size() * sizeof(TileHdr)) == 0 && T1.hnodes == T.hnodes,
          "%s: bytes depend on the thread count", name);
    std::printf("%s: %lld tiles, halo %lld\n", name, (long long)T.ntiles, (long long)T.nhalo);
}

static TileShape shape_of(int npe, int rbytes, bool grad, int max_nodes, int kib) {
    TileShape s;
    s.npe = npe;
    s.nqp = npe == 8 ? 8 : 4;
    s.rbytes = rbytes;
    s.grad = grad;
    s.max_nodes = max_nodes;
    s.max_halo = 255;
    s.threads = 1024;
    s.budget = (size_t)kib * 1024;
    return s;
}

int main() {
    const TileHostMesh a = h8(11, 9, 7, 1.0, 0.15, 3), slab = h8(12, 12, 1, 0.001, 0.0, 0), one = h8(1, 1, 1, 1.0, 0, 0);
    const TileHostMesh b = t10(6, 5);
    check_mesh("h8 11x9x7 fp64", a, shape_of(8, 8, true, 64, 160));
    check_mesh("h8 11x9x7 mixed", a, shape_of(8, 4, true, 64, 160));
    check_mesh("h8 11x9x7 dot", a, shape_of(8, 8, false, 32, 160));
    check_mesh("h8 11x9x7 fp64 16 nodes 40 KiB", a, shape_of(8, 8, true, 16, 40));
    check_mesh("h8 11x9x7 fp64 128 nodes", a, shape_of(8, 8, true, 128, 160));
    check_mesh("h8 slab", slab, shape_of(8, 8, true, 64, 160));
    check_mesh("h8 one element", one, shape_of(8, 8, true, 64, 160));
    const TileHostMesh region = h8_region(9, 7, 5);
    check_mesh("h8 region, nodes of no element", region, shape_of(8, 8, true, 64, 160));
    check_mesh("h8 region mixed", region, shape_of(8, 4, true, 64, 160));
    check_mesh("t10 6^3 fp64", b, shape_of(10, 8, true, 64, 160));
    check_mesh("t10 6^3 mixed", b, shape_of(10, 4, true, 64, 160));
    check_mesh("t10 6^3 dot", b, shape_of(10, 8, false, 32, 160));
    if (failures) return 1;
    std::printf("ok\n");
    return 0;
}
