// Block-mesh generators (input generation for the assembly path).
//   H8blockx   src/MeshHexahedronModule.jl:59-112: nodes x fastest; cells with
//              the z index innermost; local order f, f+1, f+nx+2, f+nx+1, +layer
//   T4blockx   src/MeshTetrahedronModule.jl:81-151, orientation :a (six tets
//              per cell, the same table for both cell parities, :94-96)
//   T4toT10    src/MeshTetrahedronModule.jl:158-212: local edges (1,2) (2,3)
//              (3,1) (4,1) (4,2) (4,3); a new node per edge, numbered in the
//              order the edges are first met (element, then local edge);
//              mid-edge node at the mean of the two end points.
// The first-encounter numbering is computed by sorting the edge occurrences
// (rather than by a search container): every distinct edge takes the rank of
// its earliest occurrence, which is the same numbering.
// Grid coordinates are L*i/n (exact i/n on the unit interval, the values the
// reference's linearspace produces for [0, 1]).
#include <algorithm>
#include <numeric>

#include "common.hpp"

namespace feasm {

namespace {

uint64_t splitmix(uint64_t seed, uint64_t ctr) {
    uint64_t z = seed + ctr * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void grid_nodes(double L, double W, double H, int64_t nL, int64_t nW, int64_t nH, double *xyz) {
    int64_t f = 0;
    for (int64_t k = 0; k <= nH; k++)
        for (int64_t j = 0; j <= nW; j++)
            for (int64_t i = 0; i <= nL; i++, f++) {
                xyz[3 * f + 0] = L * (double)i / (double)nL;
                xyz[3 * f + 1] = W * (double)j / (double)nW;
                xyz[3 * f + 2] = H * (double)k / (double)nH;
            }
}

// 1-based node number of the lower corner of cell (i, j, k), 0-based indices
inline int64_t cell_first(int64_t i, int64_t j, int64_t k, int64_t nL, int64_t nW) {
    return k * (nL + 1) * (nW + 1) + j * (nL + 1) + i + 1;
}

}  // namespace

void perturb_nodes(int64_t nnodes, double *xyz, const double lo[3], const double hi[3], double amp, uint64_t seed) {
    for (int64_t n = 0; n < nnodes; n++) {
        double *x = xyz + 3 * n;
        if (!(x[0] > lo[0] && x[0] < hi[0] && x[1] > lo[1] && x[1] < hi[1] && x[2] > lo[2] && x[2] < hi[2]))
            continue;
        for (int c = 0; c < 3; c++) {
            const uint64_t z = splitmix(seed, 3 * (uint64_t)n + (uint64_t)c + 1);
            const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
            x[c] = x[c] + u * amp;
        }
    }
}

void gen_h8block(double L, double W, double H, int64_t nL, int64_t nW, int64_t nH, int64_t *conn, double *xyz) {
    grid_nodes(L, W, H, nL, nW, nH, xyz);
    const int64_t layer = (nL + 1) * (nW + 1);
    int64_t e = 0;
    for (int64_t i = 0; i < nL; i++)
        for (int64_t j = 0; j < nW; j++)
            for (int64_t k = 0; k < nH; k++, e++) {
                const int64_t f = cell_first(i, j, k, nL, nW);
                const int64_t q[4] = {f, f + 1, f + (nL + 1) + 1, f + (nL + 1)};
                for (int a = 0; a < 4; a++) {
                    conn[8 * e + a] = q[a];
                    conn[8 * e + 4 + a] = q[a] + layer;
                }
            }
}

// returns the T10 node count; conn10 nelem x 10 (1-based), xyz10 row-major
int64_t gen_t10block(double L, double W, double H, int64_t nL, int64_t nW, int64_t nH, double amp_v, double amp_m,
                     uint64_t seed, int64_t *conn10, std::vector<double> &xyz10) {
    static const int kTet[6][4] = {{1, 8, 5, 6}, {3, 4, 2, 7}, {7, 2, 6, 8}, {4, 7, 8, 2}, {2, 1, 6, 8}, {4, 8, 1, 2}};
    static const int kEdge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {3, 0}, {3, 1}, {3, 2}};
    const int64_t nv = (nL + 1) * (nW + 1) * (nH + 1);
    const int64_t ne = 6 * nL * nW * nH;
    const int64_t layer = (nL + 1) * (nW + 1);
    std::vector<double> xv(3 * nv);
    grid_nodes(L, W, H, nL, nW, nH, xv.data());
    const double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {L, W, H};
    if (amp_v != 0.0) perturb_nodes(nv, xv.data(), lo, hi, amp_v, seed);
    int64_t e = 0;
    for (int64_t i = 0; i < nL; i++)
        for (int64_t j = 0; j < nW; j++)
            for (int64_t k = 0; k < nH; k++) {
                const int64_t f = cell_first(i, j, k, nL, nW);
                const int64_t nn[8] = {f, f + 1, f + (nL + 1) + 1, f + (nL + 1), 0, 0, 0, 0};
                int64_t c8[8];
                for (int a = 0; a < 4; a++) {
                    c8[a] = nn[a];
                    c8[4 + a] = nn[a] + layer;
                }
                for (int t = 0; t < 6; t++, e++)
                    for (int a = 0; a < 4; a++) conn10[10 * e + a] = c8[kTet[t][a] - 1];
            }
    // edge occurrences: key (lo, hi), occurrence number e*6 + J
    struct Occ {
        int64_t lo, hi, occ;
    };
    std::vector<Occ> occ(6 * ne);
    for (int64_t x = 0; x < ne; x++)
        for (int J = 0; J < 6; J++) {
            const int64_t a = conn10[10 * x + kEdge[J][0]], b = conn10[10 * x + kEdge[J][1]];
            occ[6 * x + J] = {std::min(a, b), std::max(a, b), 6 * x + J};
        }
    std::vector<int64_t> order(occ.size());
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int64_t p, int64_t q) {
        if (occ[p].lo != occ[q].lo) return occ[p].lo < occ[q].lo;
        if (occ[p].hi != occ[q].hi) return occ[p].hi < occ[q].hi;
        return occ[p].occ < occ[q].occ;
    });
    // group -> first occurrence; rank groups by first occurrence
    std::vector<int64_t> group_of(occ.size()), first;
    for (size_t s = 0; s < order.size(); s++) {
        const Occ &o = occ[order[s]];
        if (s == 0 || o.lo != occ[order[s - 1]].lo || o.hi != occ[order[s - 1]].hi) first.push_back(o.occ);
        group_of[order[s]] = (int64_t)first.size() - 1;
    }
    std::vector<int64_t> gorder(first.size());
    std::iota(gorder.begin(), gorder.end(), 0);
    std::sort(gorder.begin(), gorder.end(), [&](int64_t p, int64_t q) { return first[p] < first[q]; });
    std::vector<int64_t> number(first.size());
    for (size_t r = 0; r < gorder.size(); r++) number[gorder[r]] = nv + 1 + (int64_t)r;
    const int64_t ntot = nv + (int64_t)first.size();
    xyz10.assign(3 * ntot, 0.0);
    std::copy(xv.begin(), xv.end(), xyz10.begin());
    for (int64_t x = 0; x < ne; x++)
        for (int J = 0; J < 6; J++) {
            const int64_t nid = number[group_of[6 * x + J]];
            conn10[10 * x + 4 + J] = nid;
            const Occ &o = occ[6 * x + J];
            for (int c = 0; c < 3; c++)
                xyz10[3 * (nid - 1) + c] = (xyz10[3 * (o.hi - 1) + c] + xyz10[3 * (o.lo - 1) + c]) / 2;
        }
    if (amp_m != 0.0) perturb_nodes(ntot - nv, xyz10.data() + 3 * nv, lo, hi, amp_m, seed + 1);
    return ntot;
}

}  // namespace feasm
