// Quadrature rules and basis-function tables (host side, uploaded once per
// context).  The numeric literals are the ones the reference types:
//   GaussRule  src/IntegRuleModule.jl:206-398 (1D points/weights :209-216,
//              3D tensor product with the zeta index fastest :374-390)
//   TetRule    src/IntegRuleModule.jl:483-516 (8-digit 4-point literals :488-493)
//   H8 basis   src/FESetModule.jl:955-992 (products of 1 -/+ coordinate, /8)
//   T10 basis  src/FESetModule.jl:1395-1427
#include "common.hpp"

namespace feasm {

namespace {

bool gauss_1d(int order, std::vector<double> &p, std::vector<double> &w) {
    switch (order) {
        case 1: p = {0.0}; w = {2.0}; return true;
        case 2: p = {-0.577350269189626, 0.577350269189626}; w = {1.0, 1.0}; return true;
        case 3:
            p = {-0.774596669241483, 0.0, 0.774596669241483};
            w = {0.5555555555555556, 0.8888888888888889, 0.5555555555555556};
            return true;
        case 4:
            p = {-0.86113631159405, -0.33998104358486, 0.33998104358486, 0.86113631159405};
            w = {0.34785484513745, 0.65214515486255, 0.65214515486255, 0.34785484513745};
            return true;
        default: return false;
    }
}

// parametric points (nqp x 3) and weights of the rule
void rule_points(int rule, int order, std::vector<double> &pc, std::vector<double> &w) {
    pc.clear();
    w.clear();
    if (rule == FEASM_RULE_GAUSS) {
        std::vector<double> p1, w1;
        FEASM_REQUIRE(gauss_1d(order, p1, w1), FEASM_EINVAL,
                      "Gauss rule of order " + std::to_string(order) + " not available");
        const int n = order;
        for (int r = 0; r < n * n * n; r++) {
            const int i = r / (n * n), j = (r / n) % n, k = r % n;  // k fastest
            pc.insert(pc.end(), {p1[i], p1[j], p1[k]});
            w.push_back(w1[i] * w1[j] * w1[k]);
        }
        return;
    }
    FEASM_REQUIRE(rule == FEASM_RULE_TET, FEASM_EINVAL, "unknown integration rule");
    if (order == 1) {
        pc = {0.25, 0.25, 0.25};
        w = {1.0 / 6.0};
    } else if (order == 4) {
        const double lo = 0.13819660, hi = 0.58541020;
        pc = {lo, lo, lo, hi, lo, lo, lo, hi, lo, lo, lo, hi};
        w.assign(4, 0.041666666666666666667);
    } else if (order == 5) {
        const double a = 1.0 / 6.0, b = 0.25, c = 0.5;
        pc = {b, b, b, c, a, a, a, c, a, a, a, c, a, a, a};
        w = {-0.8 / 6, 0.45 / 6, 0.45 / 6, 0.45 / 6, 0.45 / 6};
    } else {
        throw Error(FEASM_EINVAL, "Unknown number of integration points");
    }
}

// H8 corner signs in the reference's local node order
const int kH8Sign[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                           {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

void h8_basis(const double *x, double *N, double *dN) {
    for (int a = 0; a < 8; a++) {
        const double f0 = 1.0 + kH8Sign[a][0] * x[0];
        const double f1 = 1.0 + kH8Sign[a][1] * x[1];
        const double f2 = 1.0 + kH8Sign[a][2] * x[2];
        N[a] = f0 * f1 * f2 / 8.0;
        dN[3 * a + 0] = kH8Sign[a][0] * (f1 * f2) / 8.0;
        dN[3 * a + 1] = kH8Sign[a][1] * (f0 * f2) / 8.0;
        dN[3 * a + 2] = kH8Sign[a][2] * (f0 * f1) / 8.0;
    }
}

void t10_basis(const double *x, double *N, double *dN) {
    const double r = x[0], s = x[1], t = x[2];
    const double m = 1 - r - s - t;
    const double Nv[10] = {m * (2 * m - 1), r * (2 * r - 1), s * (2 * s - 1), t * (2 * t - 1), 4 * m * r,
                           4 * r * s,       4 * s * m,       4 * m * t,       4 * r * t,       4 * s * t};
    const double c = -3 + 4 * r + 4 * s + 4 * t;
    const double dr[10] = {c, 4 * r - 1, 0, 0, -8 * r + 4 - 4 * s - 4 * t, 4 * s, -4 * s, -4 * t, 4 * t, 0};
    const double ds[10] = {c, 0, 4 * s - 1, 0, -4 * r, 4 * r, 4 - 4 * r - 8 * s - 4 * t, -4 * t, 0, 4 * t};
    const double dt[10] = {c, 0, 0, 4 * t - 1, -4 * r, 0, -4 * s, -8 * t + 4 - 4 * r - 4 * s, 4 * r, 4 * s};
    for (int a = 0; a < 10; a++) {
        N[a] = Nv[a];
        dN[3 * a + 0] = dr[a];
        dN[3 * a + 1] = ds[a];
        dN[3 * a + 2] = dt[a];
    }
}

}  // namespace

QTable make_qtable(int etype, int rule, int order) {
    FEASM_REQUIRE(etype == FEASM_H8 || etype == FEASM_T10, FEASM_EINVAL, "unknown element type");
    QTable t;
    t.etype = etype;
    t.rule = rule;
    t.order = order;
    t.npe = (etype == FEASM_H8) ? 8 : 10;
    std::vector<double> pc;
    rule_points(rule, order, pc, t.w);
    t.nqp = (int)t.w.size();
    t.N.resize(t.nqp * t.npe);
    t.dN.resize(t.nqp * t.npe * 3);
    for (int q = 0; q < t.nqp; q++) {
        if (etype == FEASM_H8)
            h8_basis(&pc[3 * q], &t.N[q * t.npe], &t.dN[q * t.npe * 3]);
        else
            t10_basis(&pc[3 * q], &t.N[q * t.npe], &t.dN[q * t.npe * 3]);
    }
    return t;
}

const DevQTable &device_qtable(feasm_ctx *ctx, int etype, int rule, int order) {
    const int key = etype * 1000 + rule * 100 + order;
    auto it = ctx->qtables.find(key);
    if (it != ctx->qtables.end()) return it->second;
    QTable t = make_qtable(etype, rule, order);
    DevQTable d;
    d.npe = t.npe;
    d.nqp = t.nqp;
    std::vector<double> host;
    host.insert(host.end(), t.N.begin(), t.N.end());
    host.insert(host.end(), t.dN.begin(), t.dN.end());
    host.insert(host.end(), t.w.begin(), t.w.end());
    d.buf = dev_alloc(host.size() * sizeof(double));
    FEASM_HIP(hipMemcpy(d.buf->p, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice));
    return ctx->qtables.emplace(key, d).first->second;
}

}  // namespace feasm
