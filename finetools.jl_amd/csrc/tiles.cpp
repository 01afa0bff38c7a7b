// Tiles for the fused numeric kernel (symbolic phase, host side).
//
// A tile is a set of "owned" column nodes, contiguous in a Morton (Z-curve)
// order of the node coordinates so that it is spatially compact, together
// with its halo: every element incident to an owned node, in ascending
// element number.  One workgroup integrates the halo elements into LDS and
// produces every CSC node block of the owned columns.
//
// The work of a tile is listed as block tasks: one per node block (m, c) of
// an owned column c, carrying the elements that hold both nodes in ascending
// element number (the order in which the reference's sparse() folds the
// duplicates of its element-major COO, src/AssemblyModule.jl:319-325), each
// with the local indices of the two nodes in the element.  Tasks are sorted
// by their element count, so that the lanes of one wavefront run loops of the
// same length, and the wavefronts are dealt to the SIMDs so that each gets a
// similar share of the work.
//
// Tiles are boxes of the Morton order, split in two along the highest
// varying code bit until each fits the LDS budget (fused.hpp: box_fits,
// box_split).  This host builder and the device one (tiles_gpu.hip) follow
// the same rules and produce the same bytes; the assembly uses the device
// one, and this one is the reference it is checked against.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <thread>
#include <utility>

#include "common.hpp"
#include "fused.hpp"

namespace feasm {

namespace {

using HostMesh = TileHostMesh;

// Node order along the Z-curve (see morton_code), ties by node number.
std::vector<std::pair<uint64_t, int32_t>> morton_order(const HostMesh &h) {
    const int64_t nn = h.nn;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t n = 0; n < nn; n++)
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], h.xyz[3 * n + k]);
            hi[k] = std::max(hi[k], h.xyz[3 * n + k]);
        }
    const double step = morton_step(lo, hi, nn);
    std::vector<std::pair<uint64_t, int32_t>> key(nn);
    for (int64_t n = 0; n < nn; n++) key[n] = {morton_code(&h.xyz[3 * n], lo, step), (int32_t)n};
    // stable LSD radix sort over the codes' significant bits (11 bits a
    // pass): the order of std::sort on (code, node), nodes ascending within
    // equal codes
    uint64_t maxc = 0;
    for (const auto &k : key) maxc = std::max(maxc, k.first);
    const int bits = maxc ? 64 - __builtin_clzll(maxc) : 0;
    std::vector<std::pair<uint64_t, int32_t>> tmp(nn);
    std::vector<int64_t> cnt(2049);
    for (int shift = 0; shift < bits; shift += 11) {
        std::fill(cnt.begin(), cnt.end(), 0);
        for (const auto &k : key) cnt[((k.first >> shift) & 2047u) + 1]++;
        for (int d = 0; d < 2048; d++) cnt[d + 1] += cnt[d];
        for (const auto &k : key) tmp[cnt[(k.first >> shift) & 2047u]++] = k;
        key.swap(tmp);
    }
    return key;
}

// Everything one tile contributes to the tile arrays.
struct TileOut {
    int P = 0, H = 0, NH = 0, K = 0;
    std::vector<int64_t> cbase;
    std::vector<int32_t> cdeg, hnodes;
    std::vector<uint16_t> lconn, item, toff, tpair;
    std::vector<uint32_t> task;
};

// Per-thread lookup tables of make_tile: an element's halo index, a node's
// index among the halo nodes, a row node's position in the current column
// (written for the current tile or column before they are read), a node's
// index among the tile's owned nodes (-1 when not owned; reset after each
// tile), and the tile in which an element or a node was last collected.
struct TileScratch {
    std::vector<int32_t> ehalo, npos, rpos, opos, estamp, nstamp;
    int32_t epoch = 0;
    // reused by spread_halo_banks
    std::vector<int32_t> set_ptr, occ_ptr, occ, fill, seen;
    std::vector<uint16_t> reads, lconn;
    std::vector<uint8_t> load;
    std::vector<int> order, perm;
    std::vector<std::vector<int>> slots;
    TileScratch(int64_t ne, int64_t nn)
        : ehalo(ne), npos(nn), rpos(nn), opos(nn, -1), estamp(ne, -1), nstamp(nn, -1) {}
};

// Halo numbering for the LDS banks.  Phase 2 reads 16-byte chunks of the
// halo elements' gradient records; a quarter-wavefront (16 lanes) reads its
// 16 chunks in one cycle when they fall into distinct 16-byte bank groups,
// and the group of (halo element h, node a) is (uchunks * h + a + a term
// the same for every lane) mod 16.  The records can sit in any order (every
// sum runs in global element order), so the halo elements are renumbered:
// each gets a residue class, greedily, the one that adds the fewest
// collisions with the reads of the elements placed before it (most-read
// elements first), within the number of slots each class has.  The values
// are bitwise the same; only the LDS addresses change.  Measured: C2 (H8,
// FP64) 1.21 -> 1.17 ms; no gain for T10 or mixed precision, which keep the
// ascending order (and skip the host work).
void spread_halo_banks(const TileShape &shape, int npe, TileScratch &sc, TileOut &o) {
    const int H = o.H;
    if (!shape.grad || shape.rbytes != 8 || npe != 8 || H <= 1) return;
    constexpr int W = 16;
    // read sets: per quarter-wavefront of the schedule and item step, the
    // distinct (h, node) reads of the column side and of the row side
    std::vector<int32_t> &set_ptr = sc.set_ptr, &seen = sc.seen;
    std::vector<uint16_t> &reads = sc.reads;  // h << 4 | node
    set_ptr.assign(1, 0);
    reads.clear();
    seen.assign(4096, -1);  // the set a read was last added to
    for (size_t g0 = 0; g0 < o.task.size(); g0 += W) {
        const size_t g1 = std::min(o.task.size(), g0 + W);
        int kmax = 0;
        for (size_t t = g0; t < g1; t++)
            if (o.task[t] != kNullTask) kmax = std::max(kmax, (int)((o.task[t] >> 16) & 255u));
        for (int j = 0; j < kmax; j++)
            for (int side = 0; side < 2; side++) {
                const size_t b = reads.size();
                for (size_t t = g0; t < g1; t++) {
                    if (o.task[t] == kNullTask || j >= (int)((o.task[t] >> 16) & 255u)) continue;
                    const uint16_t it = o.item[o.toff[t] + j];
                    const uint16_t r = (uint16_t)(((it >> 8) << 4) | (side ? (it & 15u) : ((it >> 4) & 15u)));
                    if (seen[r] != (int32_t)set_ptr.size()) {
                        seen[r] = (int32_t)set_ptr.size();
                        reads.push_back(r);
                    }
                }
                if (reads.size() > b) set_ptr.push_back((int32_t)reads.size());
            }
    }
    const int nsets = (int)set_ptr.size() - 1;
    // each element's reads: (set, node)
    std::vector<int32_t> &occ_ptr = sc.occ_ptr, &occ = sc.occ, &fill = sc.fill;
    occ_ptr.assign(H + 1, 0);
    for (uint16_t r : reads) occ_ptr[(r >> 4) + 1]++;
    for (int e = 0; e < H; e++) occ_ptr[e + 1] += occ_ptr[e];
    occ.resize(reads.size());
    fill.assign(occ_ptr.begin(), occ_ptr.end() - 1);
    for (int si = 0; si < nsets; si++)
        for (int32_t i = set_ptr[si]; i < set_ptr[si + 1]; i++) occ[fill[reads[i] >> 4]++] = si * 16 + (reads[i] & 15);
    // slots per residue class: slot s sits in class (uchunks * s) mod 16
    const int uc = shape.uchunks();
    std::vector<std::vector<int>> &slots = sc.slots;
    slots.resize(16);
    for (auto &v : slots) v.clear();
    for (int sl = H - 1; sl >= 0; sl--) slots[(uc * sl) & 15].push_back(sl);  // popped ascending
    std::vector<int> &order = sc.order;
    order.resize(H);
    for (int e = 0; e < H; e++) order[e] = e;
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return occ_ptr[x + 1] - occ_ptr[x] > occ_ptr[y + 1] - occ_ptr[y]; });
    // load of each set's 16 bank groups, stored twice in a row, so that the
    // loads seen by class c for node a are row[a .. a + 16) (c = 0..15)
    std::vector<uint8_t> &load = sc.load;
    load.assign((size_t)nsets * 32, 0);
    std::vector<int> &perm = sc.perm;
    perm.resize(H);
    for (int e : order) {
        uint16_t cost[16] = {};
        for (int32_t i = occ_ptr[e]; i < occ_ptr[e + 1]; i++) {
            const uint8_t *row = load.data() + (size_t)(occ[i] >> 4) * 32 + (occ[i] & 15);
            for (int c = 0; c < 16; c++) cost[c] = (uint16_t)(cost[c] + row[c]);
        }
        int best = -1;
        for (int c = 0; c < 16; c++) {
            if (slots[c].empty()) continue;
            if (best < 0 || cost[c] < cost[best] || (cost[c] == cost[best] && slots[c].size() > slots[best].size()))
                best = c;
        }
        perm[e] = slots[best].back();
        slots[best].pop_back();
        for (int32_t i = occ_ptr[e]; i < occ_ptr[e + 1]; i++) {
            uint8_t *row = load.data() + (size_t)(occ[i] >> 4) * 32;
            const int bank = (best + (occ[i] & 15)) & 15;
            if (row[bank] < 255) {
                row[bank]++;
                row[bank + 16]++;
            }
        }
    }
    for (uint16_t &it : o.item) it = (uint16_t)((perm[it >> 8] << 8) | (it & 255u));
    std::vector<uint16_t> &lc = sc.lconn;
    lc.resize(o.lconn.size());
    for (int e = 0; e < H; e++)
        for (int a = 0; a < npe; a++) lc[(size_t)perm[e] * npe + a] = o.lconn[(size_t)e * npe + a];
    std::copy(lc.begin(), lc.end(), o.lconn.begin());
}

void make_tile(const HostMesh &h, const int32_t *owned, int P, const TileShape &shape, TileScratch &sc,
               TileOut &o) {
    const int npe = h.npe;
    // the halo elements and their nodes, each once, ascending
    const int32_t epoch = ++sc.epoch;
    std::vector<int32_t> halo;
    for (int c = 0; c < P; c++)
        for (int64_t s = h.n2e_ptr[owned[c]]; s < h.n2e_ptr[owned[c] + 1]; s++) {
            const int32_t e = (int32_t)(h.n2e[s] >> 4);
            if (sc.estamp[e] != epoch) {
                sc.estamp[e] = epoch;
                halo.push_back(e);
            }
        }
    std::sort(halo.begin(), halo.end());
    std::vector<int32_t> hn;
    for (int32_t e : halo)
        for (int a = 0; a < npe; a++) {
            const int32_t v = h.conn[(int64_t)e * npe + a];
            if (sc.nstamp[v] != epoch) {
                sc.nstamp[v] = epoch;
                hn.push_back(v);
            }
        }
    std::sort(hn.begin(), hn.end());
    o.P = P;
    o.H = (int)halo.size();
    o.NH = (int)hn.size();
    FEASM_REQUIRE(o.H <= 255 && o.NH <= 65535 && npe <= 15 && P <= 255, FEASM_EUNSUPPORTED,
                  "tile exceeds the fused kernel's index widths");
    o.hnodes = hn;
    for (int i = 0; i < o.H; i++) sc.ehalo[halo[i]] = i;
    for (int i = 0; i < o.NH; i++) sc.npos[hn[i]] = i;
    o.lconn.resize((size_t)o.H * npe);
    for (int i = 0; i < o.H; i++)
        for (int a = 0; a < npe; a++) o.lconn[(size_t)i * npe + a] = (uint16_t)sc.npos[h.conn[(int64_t)halo[i] * npe + a]];
    // block tasks: items of (column c, row position r) in ascending element,
    // held in one array (task i's items at [off, off + k), tasks in (c, r)
    // order); per column, the column's items are gathered in element order,
    // counted per row, then placed
    struct Task {
        int c, r;
        bool sw, diag, sym = false;
        int off = 0, k = 0;   // its items in `its`
        int c2 = -1, r2 = 0;  // the transposed block, when its column is owned too
        bool drop = false;    // computed by its transposed block's task
    };
    std::vector<Task> tasks;
    std::vector<uint16_t> its;
    std::vector<size_t> tfirst(P);
    o.cbase.resize(P);
    o.cdeg.resize(P);
    {
        int64_t nb = 0, ni = 0;
        for (int c = 0; c < P; c++) {
            nb += h.adj_ptr[owned[c] + 1] - h.adj_ptr[owned[c]];
            ni += (h.n2e_ptr[owned[c] + 1] - h.n2e_ptr[owned[c]]) * npe;
        }
        tasks.reserve(nb);
        its.reserve(ni);
    }
    std::vector<std::pair<int, uint16_t>> colits;  // (row position, item) of the current column
    for (int c = 0; c < P; c++) {
        const int32_t node = owned[c];
        const int64_t b0 = h.adj_ptr[node], deg = h.adj_ptr[node + 1] - b0;
        const int32_t *rows = h.adj.data() + b0;
        FEASM_REQUIRE(deg <= 255, FEASM_EUNSUPPORTED, "a node has more than 255 neighbours");
        o.cbase[c] = b0;
        o.cdeg[c] = (int32_t)deg;
        const size_t tb = tasks.size();
        tfirst[c] = tb;
        for (int r = 0; r < deg; r++) {
            Task T;
            T.c = c;
            T.r = r;
            T.sw = rows[r] > node;
            T.diag = rows[r] == node;
            tasks.push_back(T);
            sc.rpos[rows[r]] = r;
        }
        colits.clear();
        for (int64_t s = h.n2e_ptr[node]; s < h.n2e_ptr[node + 1]; s++) {
            const int32_t e = (int32_t)(h.n2e[s] >> 4);
            const int ac = (int)(h.n2e[s] & 15u);
            const int hl = sc.ehalo[e];
            for (int a = 0; a < npe; a++) {
                const int32_t v = h.conn[(int64_t)e * npe + a];
                const int r = sc.rpos[v];
                FEASM_REQUIRE(r >= 0 && r < deg && rows[r] == v, FEASM_ESTATE, "pattern misses an element node pair");
                colits.push_back({r, (uint16_t)((hl << 8) | (ac << 4) | a)});
                tasks[tb + r].k++;
            }
        }
        int off = (int)its.size();
        for (int r = 0; r < deg; r++) {
            Task &T = tasks[tb + r];
            T.off = off;
            off += T.k;
            T.k = 0;
        }
        its.resize(off);
        for (const auto &ri : colits) {
            Task &T = tasks[tb + ri.first];
            its[T.off + T.k++] = ri.second;
        }
    }
    // A block whose row node is owned too has its transpose in this tile:
    // one task computes both (the element lists are the same, and so are the
    // products, in the same order).
    {
        for (int c = 0; c < P; c++) sc.opos[owned[c]] = c;
        for (Task &T : tasks) {
            if (T.diag || T.drop) continue;
            const int32_t m = h.adj[o.cbase[T.c] + T.r];
            const int c2 = sc.opos[m];
            if (c2 < T.c) continue;  // not owned (-1), or its task is the one kept
            const int32_t *rows2 = h.adj.data() + o.cbase[c2];
            const int r2 = (int)(std::lower_bound(rows2, rows2 + o.cdeg[c2], owned[T.c]) - rows2);
            FEASM_REQUIRE(r2 < o.cdeg[c2] && rows2[r2] == owned[T.c], FEASM_ESTATE, "pattern is not symmetric");
            T.c2 = c2;
            T.r2 = r2;
            tasks[tfirst[c2] + r2].drop = true;
        }
        for (int c = 0; c < P; c++) sc.opos[owned[c]] = -1;
        tasks.erase(std::remove_if(tasks.begin(), tasks.end(), [](const Task &T) { return T.drop; }), tasks.end());
    }
    for (const Task &T : tasks)  // the kernel reads a task's first item unconditionally
        FEASM_REQUIRE(T.k > 0, FEASM_ESTATE, "a pattern block has no element");
    // A diagonal block whose items all pair a node with itself (no element
    // repeats the node) is a sum of symmetric products: the kernel forms its
    // upper triangle from one vector per point.
    for (Task &T : tasks) {
        T.sym = T.diag;
        for (int j = 0; T.sym && j < T.k; j++)
            if (((its[T.off + j] >> 4) & 15u) != (its[T.off + j] & 15u)) T.sym = false;
    }
    // longest first; among equal counts the symmetric diagonal blocks first,
    // so that they fill wavefronts of their own (a stable counting sort on
    // (k, sym): the order of std::stable_sort by that key)
    int K = 0;
    for (const Task &T : tasks) K = std::max(K, T.k);
    FEASM_REQUIRE(K <= 255, FEASM_EUNSUPPORTED, "a node block is shared by more than 255 elements");
    {
        auto bucket = [K](const Task &T) { return 2 * (K - T.k) + (T.sym ? 0 : 1); };
        std::vector<int> start(2 * K + 3, 0);
        for (const Task &T : tasks) start[bucket(T) + 1]++;
        for (size_t b = 1; b < start.size(); b++) start[b] += start[b - 1];
        std::vector<Task> sorted(tasks.size());
        for (const Task &T : tasks) sorted[start[bucket(T)]++] = T;
        tasks.swap(sorted);
    }
    o.K = K;
    // Wavefront schedule: 64 consecutive tasks run as one wavefront, and the
    // wavefronts of a workgroup go to the 4 SIMDs in turn.  Chunks of 64
    // tasks of equal element count (costed by that count plus the material
    // and store work) are dealt to the SIMDs longest first, each to the least
    // loaded one; chunk slots a SIMD does not use are filled with null tasks.
    constexpr int W = 64, NS = 4;
    std::vector<std::vector<int>> queue(NS);
    {
        std::vector<double> load(NS, 0.0);
        for (size_t c0 = 0; c0 < tasks.size(); c0 += W) {
            const double cost = (double)tasks[c0].k + 2.0;
            const int q = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            queue[q].push_back((int)c0);
            load[q] += cost;
        }
    }
    size_t rounds = 0;
    for (auto &q : queue) rounds = std::max(rounds, q.size());
    if (rounds * NS * W > tasks.size() + kPadTasks) {  // too much padding: plain order
        for (auto &q : queue) q.clear();
        int i = 0;
        for (size_t c0 = 0; c0 < tasks.size(); c0 += W, i++) queue[i % NS].push_back((int)c0);
        rounds = 0;
        for (auto &q : queue) rounds = std::max(rounds, q.size());
    }
    o.task.clear();
    o.toff.clear();
    o.tpair.clear();
    o.item.clear();
    o.task.reserve(rounds * NS * W);
    o.toff.reserve(rounds * NS * W);
    o.tpair.reserve(rounds * NS * W);
    o.item.reserve(its.size());
    for (size_t rr = 0; rr < rounds; rr++)
        for (int q = 0; q < NS; q++) {
            const int c0 = rr < queue[q].size() ? queue[q][rr] : -1;
            for (int l = 0; l < W; l++) {
                const size_t t = (size_t)c0 + l;
                if (c0 < 0 || t >= tasks.size()) {
                    o.task.push_back(kNullTask);
                    o.toff.push_back(0);
                    o.tpair.push_back(kNoPair);
                    continue;
                }
                const Task &T = tasks[t];
                FEASM_REQUIRE(o.item.size() + (size_t)T.k <= 65535, FEASM_EUNSUPPORTED,
                              "tile holds too many (block, element) items");
                o.task.push_back((uint32_t)T.c | ((uint32_t)T.r << 8) | ((uint32_t)T.k << 16) |
                                 ((uint32_t)T.sw << 24) | ((uint32_t)T.diag << 25) |
                                 ((uint32_t)T.sym << 26));
                o.toff.push_back((uint16_t)o.item.size());
                o.tpair.push_back(T.c2 < 0 ? kNoPair : (uint16_t)(T.c2 | (T.r2 << 8)));
                o.item.insert(o.item.end(), its.begin() + T.off, its.begin() + T.off + T.k);
            }
        }
    // trailing null tasks are not needed
    while (!o.task.empty() && o.task.back() == kNullTask) {
        o.task.pop_back();
        o.toff.pop_back();
        o.tpair.pop_back();
    }
    spread_halo_banks(shape, npe, sc, o);
}

}  // namespace

namespace {
// FEASM_TILE_TIMING=1 prints the tile builder's phase times to stderr
struct PhaseClock {
    bool on = std::getenv("FEASM_TILE_TIMING") != nullptr;
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    void lap(const char *what) {
        if (!on) return;
        const auto n = std::chrono::steady_clock::now();
        std::fprintf(stderr, "tiles: %-24s %8.2f ms\n", what, std::chrono::duration<double, std::milli>(n - t).count());
        t = n;
    }
};
}  // namespace

void build_tiles_host(const TileHostMesh &h, const TileShape &shape, int max_threads, HostTiles &out) {
    PhaseClock clk;
    const auto key = morton_order(h);
    clk.lap("morton order");
    const int64_t nn = h.nn;
    const int npe = h.npe;

    // ---- partition the Morton order into tiles (fused.hpp: box_fits and
    // box_split), depth first, so that the tiles come out in Morton order
    std::vector<uint64_t> code(nn);
    for (int64_t i = 0; i < nn; i++) code[i] = key[i].first;
    std::vector<int32_t> estamp(h.ne, -1), nstamp(nn, -1);
    int32_t epoch = 0;
    auto sizes = [&](int64_t s0, int64_t s1) {  // exact sizes of the box [s0, s1)
        BoxSizes b;
        b.P = s1 - s0;
        if (b.P > box_max_nodes(shape)) return b;  // does not fit, whatever the rest
        epoch++;
        for (int64_t i = s0; i < s1; i++) {
            const int32_t n = key[i].second;
            const int64_t d = h.n2e_ptr[n + 1] - h.n2e_ptr[n];
            b.cand += d;
            b.K = std::max(b.K, d);
            b.NT += h.adj_ptr[n + 1] - h.adj_ptr[n];
            for (int64_t q = h.n2e_ptr[n]; q < h.n2e_ptr[n + 1]; q++) {
                const int32_t e = (int32_t)(h.n2e[q] >> 4);
                if (estamp[e] == epoch) continue;
                estamp[e] = epoch;
                b.H++;
                for (int a = 0; a < npe; a++) {
                    const int32_t v = h.conn[(int64_t)e * npe + a];
                    if (nstamp[v] != epoch) {
                        nstamp[v] = epoch;
                        b.NH++;
                    }
                }
            }
        }
        b.NI = b.cand * npe;
        return b;
    };
    std::vector<int64_t> starts;
    if (nn > 0) {
        std::vector<std::pair<int64_t, int64_t>> stack{{0, nn}};
        while (!stack.empty()) {
            const auto [s0, s1] = stack.back();
            stack.pop_back();
            if (box_fits(shape, sizes(s0, s1))) {
                starts.push_back(s0);
                continue;
            }
            FEASM_REQUIRE(s1 - s0 > 1, FEASM_EUNSUPPORTED, "a node has too many incident elements for the fused kernel");
            const int64_t mid = box_split(code.data(), s0, s1);
            stack.push_back({mid, s1});
            stack.push_back({s0, mid});
        }
    }
    clk.lap("partition");
    starts.push_back(nn);
    const int64_t ntiles = (int64_t)starts.size() - 1;

    // ---- per-tile arrays, built in parallel
    std::vector<int32_t> order(nn);
    for (int64_t i = 0; i < nn; i++) order[i] = key[i].second;
    std::vector<TileOut> outs(ntiles);
    const int nthreads = (int)std::max<int64_t>(
        1, std::min<int64_t>(std::max(1, max_threads),
                             std::min<int64_t>(std::thread::hardware_concurrency(), ntiles / 64 + 1)));
    std::vector<std::thread> pool;
    std::vector<std::string> errs(nthreads);
    std::vector<int> codes(nthreads, 0);
    for (int w = 0; w < nthreads; w++)
        pool.emplace_back([&, w]() {
            try {
                TileScratch sc(h.ne, nn);
                for (int64_t t = w; t < ntiles; t += nthreads)
                    make_tile(h, order.data() + starts[t], (int)(starts[t + 1] - starts[t]), shape, sc, outs[t]);
            } catch (const Error &e) {
                codes[w] = e.code;
                errs[w] = e.what();
            } catch (const std::exception &e) {
                codes[w] = FEASM_ENOMEM;
                errs[w] = e.what();
            }
        });
    for (auto &th : pool) th.join();
    for (int w = 0; w < nthreads; w++)
        if (codes[w]) throw Error(codes[w], errs[w]);
    clk.lap("per-tile arrays");

    HostTiles *T = &out;
    T->ntiles = ntiles;
    std::vector<TileHdr> &hdr = T->hdr;
    hdr.assign(ntiles, TileHdr{});
    int64_t nX = 0, nB = 0;
    for (int64_t t = 0; t < ntiles; t++) {
        const TileOut &o = outs[t];
        const TileShape::Layout L =
            shape.layout(o.H, o.P, o.NH, o.K, (int64_t)o.task.size(), (int64_t)o.item.size());
        FEASM_REQUIRE(L.total <= shape.budget, FEASM_ESTATE, "a tile exceeds the LDS budget it was sized for");
        TileHdr &d = hdr[t];
        d.P = o.P;
        d.H = o.H;
        d.NH = o.NH;
        d.NT = (int32_t)o.task.size();
        d.K = o.K;
        d.NI = (int32_t)o.item.size();
        d.x0 = (int32_t)nX;
        d.nbytes = (int32_t)(L.U - L.cbase);
        d.blob = nB;
        nX += o.NH;
        nB += d.nbytes;
        T->nhalo += o.H;
        T->ntasks += d.NT;
        T->nitems += d.NI;
        T->max_nodes = std::max(T->max_nodes, o.P);
        T->max_halo = std::max(T->max_halo, o.H);
        T->max_k = std::max(T->max_k, o.K);
    }
    FEASM_REQUIRE(nX < INT32_MAX, FEASM_EUNSUPPORTED, "mesh too large for 32-bit tile offsets");
    T->blob_bytes = nB;
    const size_t blob_size = (size_t)std::max<int64_t>(nB, 16);
    T->blob.reset(new unsigned char[blob_size]);  // every byte is written below
    T->blob_size = blob_size;
    unsigned char *const blob_base = T->blob.get();
    std::memset(blob_base + nB, 0, blob_size - (size_t)nB);
    std::vector<int32_t> &hn = T->hnodes;
    hn.assign(std::max<int64_t>(nX, 1), 0);
    auto fill = [&](int64_t t) {
        const TileOut &o = outs[t];
        const TileHdr &d = hdr[t];
        const TileShape::Layout L = shape.layout(d.H, d.P, d.NH, d.K, d.NT, d.NI);
        unsigned char *b = blob_base + d.blob;
        std::memset(b, 0, (size_t)d.nbytes);  // alignment gaps
        auto put = [&](size_t off, const void *src, size_t bytes) {
            if (bytes) std::memcpy(b + (off - L.cbase), src, bytes);
        };
        put(L.cbase, o.cbase.data(), o.cbase.size() * 8);
        put(L.cdeg, o.cdeg.data(), o.cdeg.size() * 4);
        uint64_t *rec = reinterpret_cast<uint64_t *>(b + (L.task - L.cbase));  // 16-byte aligned
        for (size_t k = 0; k < o.task.size(); k++)
            rec[k] = (uint64_t)o.task[k] | ((uint64_t)o.toff[k] << 32) | ((uint64_t)o.tpair[k] << 48);
        put(L.item, o.item.data(), o.item.size() * 2);
        put(L.lconn, o.lconn.data(), o.lconn.size() * 2);
        double *X = reinterpret_cast<double *>(b + (L.X - L.cbase));  // (x, y) pairs, then the z values
        for (int i = 0; i < o.NH; i++) {
            X[(size_t)2 * i] = h.xyz[(size_t)3 * o.hnodes[i]];
            X[(size_t)2 * i + 1] = h.xyz[(size_t)3 * o.hnodes[i] + 1];
            X[(size_t)2 * o.NH + i] = h.xyz[(size_t)3 * o.hnodes[i] + 2];
        }
        std::copy(o.hnodes.begin(), o.hnodes.end(), hn.begin() + d.x0);
    };
    {
        std::vector<std::thread> fillers;
        for (int w = 0; w < nthreads; w++)
            fillers.emplace_back([&, w]() {
                for (int64_t t = w; t < ntiles; t += nthreads) fill(t);
            });
        for (auto &th : fillers) th.join();
    }
    clk.lap("blobs");
}

double morton_step(const double lo[3], const double hi[3], int64_t nn) {
    double vol = 1, span = 0;
    int dims = 0;
    for (int k = 0; k < 3; k++) {
        span = std::max(span, hi[k] - lo[k]);
        if (hi[k] - lo[k] > 0) {
            vol *= hi[k] - lo[k];
            dims++;
        }
    }
    double step = (dims > 0 && nn > 1) ? std::pow(vol / (double)nn, 1.0 / dims) : 1.0;
    if (!(step > 0) || span / step > 2e6) step = (span > 0 ? span : 1.0) / 2e6;
    return step;
}

void tile_host_mesh(feasm_mesh *m, TileHostMesh &h) {
    const Pattern &p = *m->pattern;
    FEASM_HIP(hipStreamSynchronize(m->ctx->stream));  // mesh uploads may still be in flight on it
    h.npe = m->npe;
    h.nn = m->nnodes;
    h.ne = m->nelem;
    h.conn.resize(h.ne * h.npe);
    h.xyz.resize(h.nn * 3);
    h.n2e_ptr.resize(h.nn + 1);
    h.n2e.resize(h.ne * h.npe);
    h.adj_ptr.resize(h.nn + 1);
    FEASM_HIP(hipMemcpy(h.conn.data(), m->conn->p, h.conn.size() * 4, hipMemcpyDeviceToHost));
    FEASM_HIP(hipMemcpy(h.xyz.data(), m->xyz->p, h.xyz.size() * 8, hipMemcpyDeviceToHost));
    FEASM_HIP(hipMemcpy(h.n2e_ptr.data(), m->n2e_ptr->p, (h.nn + 1) * 8, hipMemcpyDeviceToHost));
    FEASM_HIP(hipMemcpy(h.n2e.data(), m->n2e->p, h.n2e.size() * 4, hipMemcpyDeviceToHost));
    FEASM_HIP(hipMemcpy(h.adj_ptr.data(), p.adj_ptr->p, (h.nn + 1) * 8, hipMemcpyDeviceToHost));
    h.adj.resize(h.adj_ptr[h.nn]);
    FEASM_HIP(hipMemcpy(h.adj.data(), p.adj->p, h.adj.size() * 4, hipMemcpyDeviceToHost));
}

int64_t compare_tile_builders(feasm_mesh *m, const TileShape &shape, int64_t *ntiles, double *device_ms) {
    FEASM_HIP(hipStreamSynchronize(m->ctx->stream));
    const auto t0 = std::chrono::steady_clock::now();
    const std::shared_ptr<TileSet> D = build_tiles_device(m, shape);
    FEASM_HIP(hipStreamSynchronize(m->ctx->stream));
    *device_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *ntiles = D->ntiles;
    TileHostMesh h;
    tile_host_mesh(m, h);
    HostTiles ht;
    build_tiles_host(h, shape, 16, ht);
    std::vector<TileHdr> dh(D->ntiles);
    std::vector<unsigned char> db(D->blob_bytes);
    std::vector<int32_t> dn(ht.hnodes.size());
    if (D->ntiles) FEASM_HIP(hipMemcpy(dh.data(), D->hdr->p, dh.size() * sizeof(TileHdr), hipMemcpyDeviceToHost));
    if (D->blob_bytes) FEASM_HIP(hipMemcpy(db.data(), D->blob->p, db.size(), hipMemcpyDeviceToHost));
    const int64_t nt = std::min(D->ntiles, ht.ntiles);
    const size_t nx = (size_t)std::min<int64_t>((int64_t)dn.size(), (int64_t)(D->hnodes->bytes / 4));
    if (nx) FEASM_HIP(hipMemcpy(dn.data(), D->hnodes->p, nx * 4, hipMemcpyDeviceToHost));
    for (int64_t t = 0; t < nt; t++) {
        const TileHdr &a = dh[t], &b = ht.hdr[t];
        if (std::memcmp(&a, &b, sizeof(TileHdr)) != 0) return t;
        if (a.blob + a.nbytes > D->blob_bytes ||
            std::memcmp(db.data() + a.blob, ht.blob.get() + b.blob, (size_t)a.nbytes) != 0)
            return t;
        if ((size_t)a.x0 + a.NH > nx ||
            std::memcmp(dn.data() + a.x0, ht.hnodes.data() + b.x0, (size_t)a.NH * 4) != 0)
            return t;
    }
    if (D->ntiles != ht.ntiles) return nt;
    if (D->nhalo != ht.nhalo || D->ntasks != ht.ntasks || D->nitems != ht.nitems || D->blob_bytes != ht.blob_bytes ||
        D->max_nodes != ht.max_nodes || D->max_halo != ht.max_halo || D->max_k != ht.max_k)
        return std::max<int64_t>(nt - 1, 0);
    return -1;
}

std::shared_ptr<TileSet> build_tiles(feasm_mesh *m, const TileShape &shape, hipStream_t s) {
    return build_tiles_device(m, shape, s);
}

}  // namespace feasm
