/*
 * feasm_oracle.c -- CPU restatement of FinEtools.jl's finite-element assembly
 * hot path (element matrices at quadrature points -> SysmatAssemblerSparse COO
 * -> SparseArrays.sparse CSC).
 *
 * THIS IS TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only the test suite, the smoke
 * check and the CPU-baseline leg of the benchmark load it, and only as the
 * checker.  The product library never links or calls anything in this file.
 *
 * Every routine follows the reference file:line given in its comment, in the
 * reference's own loop order and association order (compiled with
 * -ffp-contract=off so that no multiply-add is fused behind the reference's
 * back).  Where the reference relies on a third-party routine (Julia's
 * SparseArrays.sparse, LoopVectorization's @avx in mulCAtB!) the published
 * algorithm is restated and the fact is noted.
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   rules          literal constants + sum of weights + exact monomial integrals
 *   shape functions partition of unity, Kronecker delta at nodes, sum of grads = 0,
 *                  golden H8 gradN at (0.57,0.57,-0.57) (test/test_basics.jl:1175-1211)
 *   jacobian/gradN affine maps (exact), brute-force numpy inverse
 *   element mats   closed-form unit-cube Laplacian, rigid-body null space,
 *                  patch energies (test/test_forms.jl), mass sums to rho*V
 *   assemble!/sparse  7x7 known answers (test/test_basics.jl:78-133,1464-1613),
 *                  numpy dense brute force, duplicate order
 *   meshes         H8 node/element order closed forms, T10blockx 5x10x5 counts
 *                  (test/test_meshing.jl:1458-1475), H8 degree 8..27
 *                  (test/test_miscellaneous.jl:931-952)
 * Parity unpinned at the bit level: (1) LoopVectorization's reassociation of
 * mulCAtB! (src/MatrixUtilityModule.jl:486-492) -- restated as a plain ascending
 * sum; (2) SparseArrays.sparse's duplicate-summation order -- restated as the
 * stdlib's documented counting-sort left fold in input order.  Both are pinned
 * to the tolerances the reference's own tests use.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#define ORC_H8 0
#define ORC_T10 1

#define ORC_FORM_DIFFUSION_ISO 0
#define ORC_FORM_DIFFUSION_MAT 1
#define ORC_FORM_LIN_ELASTIC 2
#define ORC_FORM_DOT 3

/* ------------------------------------------------------------------------- */
/* Quadrature rules                                                          */
/* ------------------------------------------------------------------------- */

/* GaussRule(dim=3, order) -- src/IntegRuleModule.jl:206-398.  The 1D literals
 * are the ones typed in the source (:209-216); the tensor product loop is
 * i (xi), j (eta), k (zeta) with k fastest (:381-389); weights are the product
 * w[i]*w[j]*w[k] in that association.  Returns npts or -1. */
int orc_gauss_rule3(int order, double *pc, double *w)
{
    double p1[4], w1[4];
    if (order == 1) {
        p1[0] = 0.0; w1[0] = 2.0;
    } else if (order == 2) {
        p1[0] = -0.577350269189626; p1[1] = 0.577350269189626;
        w1[0] = 1.0; w1[1] = 1.0;
    } else if (order == 3) {
        p1[0] = -0.774596669241483; p1[1] = 0.0; p1[2] = 0.774596669241483;
        w1[0] = 0.5555555555555556; w1[1] = 0.8888888888888889; w1[2] = 0.5555555555555556;
    } else if (order == 4) {
        p1[0] = -0.86113631159405; p1[1] = -0.33998104358486;
        p1[2] = 0.33998104358486; p1[3] = 0.86113631159405;
        w1[0] = 0.34785484513745; w1[1] = 0.65214515486255;
        w1[2] = 0.65214515486255; w1[3] = 0.34785484513745;
    } else {
        return -1;
    }
    int r = 0;
    for (int i = 0; i < order; i++)
        for (int j = 0; j < order; j++)
            for (int k = 0; k < order; k++) {
                pc[3 * r + 0] = p1[i];
                pc[3 * r + 1] = p1[j];
                pc[3 * r + 2] = p1[k];
                w[r] = w1[i] * w1[j] * w1[k];
                r++;
            }
    return r;
}

/* TetRule(npts) -- src/IntegRuleModule.jl:483-516.  The four-point rule uses the
 * 8-digit literals 0.13819660 / 0.58541020 exactly as typed (:488-493); the
 * five-point rule is Zienkiewicz #3 with weights [d e e e e]/6 (:496-510). */
int orc_tet_rule(int npts, double *pc, double *w)
{
    if (npts == 1) {
        pc[0] = 0.25; pc[1] = 0.25; pc[2] = 0.25;
        w[0] = 1.0 / 6.0;
        return 1;
    }
    if (npts == 4) {
        const double a = 0.13819660, b = 0.58541020;
        const double t[12] = {a, a, a, b, a, a, a, b, a, a, a, b};
        for (int i = 0; i < 12; i++) pc[i] = t[i];
        for (int i = 0; i < 4; i++) w[i] = 0.041666666666666666667;
        return 4;
    }
    if (npts == 5) {
        const double a = 1.0 / 6.0, b = 0.25, c = 0.5, d = -0.8, e = 0.45;
        const double t[15] = {b, b, b, c, a, a, a, c, a, a, a, c, a, a, a};
        for (int i = 0; i < 15; i++) pc[i] = t[i];
        w[0] = d / 6; w[1] = e / 6; w[2] = e / 6; w[3] = e / 6; w[4] = e / 6;
        return 5;
    }
    return -1;
}

/* ------------------------------------------------------------------------- */
/* Shape functions                                                           */
/* ------------------------------------------------------------------------- */

/* FESetH8 privbfun -- src/FESetModule.jl:955-975: products, then "/ 8.0". */
void orc_h8_bfun(const double *x, double *N)
{
    double omx = (1.0 - x[0]), ome = (1.0 - x[1]), omt = (1.0 - x[2]);
    double opx = (1.0 + x[0]), ope = (1.0 + x[1]), opt = (1.0 + x[2]);
    N[0] = omx * ome * omt / 8.0;
    N[1] = opx * ome * omt / 8.0;
    N[2] = opx * ope * omt / 8.0;
    N[3] = omx * ope * omt / 8.0;
    N[4] = omx * ome * opt / 8.0;
    N[5] = opx * ome * opt / 8.0;
    N[6] = opx * ope * opt / 8.0;
    N[7] = omx * ope * opt / 8.0;
}

/* FESetH8 privbfundpar -- src/FESetModule.jl:977-992.  dN is 8x3 row-major
 * (dN[3*a+k] = dN_a / d xi_k); the matrix literal is transposed then /8.0. */
void orc_h8_bfundpar(const double *x, double *dN)
{
    double omxi = (1.0 - x[0]), ometa = (1.0 - x[1]), omth = (1.0 - x[2]);
    double opxi = (1.0 + x[0]), opeta = (1.0 + x[1]), opth = (1.0 + x[2]);
    const double r0[8] = {-ometa * omth, ometa * omth, opeta * omth, -opeta * omth,
                          -ometa * opth, ometa * opth, opeta * opth, -opeta * opth};
    const double r1[8] = {-omxi * omth, -opxi * omth, opxi * omth, omxi * omth,
                          -omxi * opth, -opxi * opth, opxi * opth, omxi * opth};
    const double r2[8] = {-omxi * ometa, -opxi * ometa, -opxi * opeta, -omxi * opeta,
                          omxi * ometa, opxi * ometa, opxi * opeta, omxi * opeta};
    for (int a = 0; a < 8; a++) {
        dN[3 * a + 0] = r0[a] / 8.0;
        dN[3 * a + 1] = r1[a] / 8.0;
        dN[3 * a + 2] = r2[a] / 8.0;
    }
}

/* FESetT10 privbfun -- src/FESetModule.jl:1395-1412 (Julia evaluates
 * 1 - r - s - t left to right). */
void orc_t10_bfun(const double *x, double *N)
{
    double r = x[0], s = x[1], t = x[2];
    double m = 1 - r - s - t;
    N[0] = m * (2 * m - 1);
    N[1] = r * (2 * r - 1);
    N[2] = s * (2 * s - 1);
    N[3] = t * (2 * t - 1);
    N[4] = 4 * m * r;
    N[5] = 4 * r * s;
    N[6] = 4 * s * m;
    N[7] = 4 * m * t;
    N[8] = 4 * r * t;
    N[9] = 4 * s * t;
}

/* FESetT10 privbfundpar -- src/FESetModule.jl:1414-1427 (rows of the literal are
 * d/dr, d/ds, d/dt; transposed to 10x3). */
void orc_t10_bfundpar(const double *x, double *dN)
{
    double r = x[0], s = x[1], t = x[2];
    double c = -3 + 4 * r + 4 * s + 4 * t;
    const double d0[10] = {c, 4 * r - 1, 0, 0, -8 * r + 4 - 4 * s - 4 * t, 4 * s, -4 * s, -4 * t, 4 * t, 0};
    const double d1[10] = {c, 0, 4 * s - 1, 0, -4 * r, 4 * r, 4 - 4 * r - 8 * s - 4 * t, -4 * t, 0, 4 * t};
    const double d2[10] = {c, 0, 0, 4 * t - 1, -4 * r, 0, -4 * s, -8 * t + 4 - 4 * r - 4 * s, 4 * r, 4 * s};
    for (int a = 0; a < 10; a++) {
        dN[3 * a + 0] = d0[a];
        dN[3 * a + 1] = d1[a];
        dN[3 * a + 2] = d2[a];
    }
}

/* integrationdata -- src/IntegDomainModule.jl:632-648: per quadrature point the
 * basis values and parametric gradients.  Returns npts; N is npts x npe,
 * dN is npts x npe x 3. rule: 0 = GaussRule(3, order), 1 = TetRule(order). */
int orc_integration_data(int etype, int rule, int order, double *N, double *dN, double *w)
{
    double pc[3 * 125];
    int npts = (rule == 0) ? orc_gauss_rule3(order, pc, w) : orc_tet_rule(order, pc, w);
    if (npts < 0) return -1;
    int npe = (etype == ORC_H8) ? 8 : 10;
    for (int q = 0; q < npts; q++) {
        if (etype == ORC_H8) {
            orc_h8_bfun(pc + 3 * q, N + q * npe);
            orc_h8_bfundpar(pc + 3 * q, dN + q * npe * 3);
        } else {
            orc_t10_bfun(pc + 3 * q, N + q * npe);
            orc_t10_bfundpar(pc + 3 * q, dN + q * npe * 3);
        }
    }
    return npts;
}

/* ------------------------------------------------------------------------- */
/* Per-quadrature-point kernels                                              */
/* ------------------------------------------------------------------------- */

/* locjac! -- src/MatrixUtilityModule.jl:60-68 via mulCAtB! (:478-522):
 * loc = N' * X (1x3), J = X' * dN (3x3, J[i][k] = sum_a X[a][i] dN[a][k]).
 * The sum runs over a ascending (the @avx reassociation is not reproduced). */
void orc_locjac(int npe, const double *X, const double *N, const double *dN, double *loc, double *J)
{
    for (int n = 0; n < 3; n++) {
        double c = 0.0;
        for (int k = 0; k < npe; k++) c += N[k] * X[3 * k + n];
        loc[n] = c;
    }
    for (int n = 0; n < 3; n++)
        for (int m = 0; m < 3; m++) {
            double c = 0.0;
            for (int k = 0; k < npe; k++) c += X[3 * k + m] * dN[3 * k + n];
            J[3 * m + n] = c;
        }
}

/* Jacobian (3-manifold) -- src/FESetModule.jl:479-489: cofactor expansion on
 * the first row; no positivity check (src/IntegDomainModule.jl:567-575). */
double orc_jacobian_det(const double *J)
{
#define Jm(i, j) J[3 * ((i)-1) + ((j)-1)]
    return (Jm(1, 1) * (Jm(2, 2) * Jm(3, 3) - Jm(3, 2) * Jm(2, 3)) -
            Jm(1, 2) * (Jm(2, 1) * Jm(3, 3) - Jm(2, 3) * Jm(3, 1)) +
            Jm(1, 3) * (Jm(2, 1) * Jm(3, 2) - Jm(2, 2) * Jm(3, 1)));
#undef Jm
}

/* gradN! (3-manifold) -- src/FESetModule.jl:507-544: adjugate times 1/det,
 * then gradN[r][c] = sum_k dN[r][k] * inv[k][c] in the order k = 1, 2, 3. */
void orc_gradN(int npe, const double *dN, const double *J, double *gradN)
{
#define R(i, j) J[3 * ((i)-1) + ((j)-1)]
    double invdet = 1.0 / (+R(1, 1) * (R(2, 2) * R(3, 3) - R(3, 2) * R(2, 3)) -
                           R(1, 2) * (R(2, 1) * R(3, 3) - R(2, 3) * R(3, 1)) +
                           R(1, 3) * (R(2, 1) * R(3, 2) - R(2, 2) * R(3, 1)));
    double i11 = (R(2, 2) * R(3, 3) - R(3, 2) * R(2, 3)) * invdet;
    double i12 = -(R(1, 2) * R(3, 3) - R(1, 3) * R(3, 2)) * invdet;
    double i13 = (R(1, 2) * R(2, 3) - R(1, 3) * R(2, 2)) * invdet;
    double i21 = -(R(2, 1) * R(3, 3) - R(2, 3) * R(3, 1)) * invdet;
    double i22 = (R(1, 1) * R(3, 3) - R(1, 3) * R(3, 1)) * invdet;
    double i23 = -(R(1, 1) * R(2, 3) - R(2, 1) * R(1, 3)) * invdet;
    double i31 = (R(2, 1) * R(3, 2) - R(3, 1) * R(2, 2)) * invdet;
    double i32 = -(R(1, 1) * R(3, 2) - R(3, 1) * R(1, 2)) * invdet;
    double i33 = (R(1, 1) * R(2, 2) - R(2, 1) * R(1, 2)) * invdet;
#undef R
    for (int r = 0; r < npe; r++) {
        const double *g = dN + 3 * r;
        gradN[3 * r + 0] = g[0] * i11 + g[1] * i21 + g[2] * i31;
        gradN[3 * r + 1] = g[0] * i12 + g[1] * i22 + g[2] * i32;
        gradN[3 * r + 2] = g[0] * i13 + g[1] * i23 + g[2] * i33;
    }
}

/* blmat!(DeforModelRed3D) -- src/DeforModelRedModule.jl:447-472, strain order
 * xx, yy, zz, xy, xz, yz (:150-160).  Rm is the material frame (identity for
 * the default CSys(3), src/CSysModule.jl:154-181).  B is 6 x 3npe row-major. */
void orc_blmat3d(int npe, const double *gradN, const double *Rm, double *B)
{
    int nc = 3 * npe;
    memset(B, 0, sizeof(double) * 6 * nc);
#define RM(i, j) Rm[3 * ((i)-1) + ((j)-1)]
#define G(i, j) gradN[3 * ((i)-1) + ((j)-1)]
    for (int i = 1; i <= npe; i++) {
        int k = 3 * (i - 1);
        for (int j = 1; j <= 3; j++) {
            int kj = k + j - 1;
            B[0 * nc + kj] = G(i, 1) * RM(j, 1);
            B[1 * nc + kj] = G(i, 2) * RM(j, 2);
            B[2 * nc + kj] = G(i, 3) * RM(j, 3);
            B[3 * nc + kj] = G(i, 2) * RM(j, 1) + G(i, 1) * RM(j, 2);
            B[4 * nc + kj] = G(i, 3) * RM(j, 1) + G(i, 1) * RM(j, 3);
            B[5 * nc + kj] = G(i, 3) * RM(j, 2) + G(i, 2) * RM(j, 3);
        }
    }
#undef RM
#undef G
}

/* add_mggt_ut_only! -- src/MatrixUtilityModule.jl:84-98.  Ke is n x n
 * column-major (Ke[mx + n*nx]); loop order nx, px, mx <= nx. */
void orc_add_mggt_ut_only(int n, double *Ke, const double *gradN, double mult)
{
    for (int nx = 0; nx < n; nx++)
        for (int px = 0; px < 3; px++) {
            double a = (mult)*gradN[3 * nx + px];
            for (int mx = 0; mx <= nx; mx++) Ke[mx + n * nx] += gradN[3 * mx + px] * a;
        }
}

/* add_gkgt_ut_only! -- src/MatrixUtilityModule.jl:120-153 (kappa 3x3 row-major). */
void orc_add_gkgt_ut_only(int n, double *Ke, const double *gradN, double Jac_w, const double *kappa)
{
    double kg[3 * 10];
    for (int nx = 0; nx < n; nx++)
        for (int mx = 0; mx < 3; mx++) {
            double acc = 0.0;
            for (int px = 0; px < 3; px++) acc += kappa[3 * mx + px] * gradN[3 * nx + px];
            kg[mx * n + nx] = Jac_w * acc;
        }
    for (int nx = 0; nx < n; nx++)
        for (int mx = 0; mx <= nx; mx++) {
            double acc = 0.0;
            for (int px = 0; px < 3; px++) acc += gradN[3 * mx + px] * kg[px * n + nx];
            Ke[mx + n * nx] += acc;
        }
}

/* add_btdb_ut_only! -- src/MatrixUtilityModule.jl:189-216.  B is nstr x Kedim
 * row-major, D nstr x nstr row-major; DB = Jac_w * (D*B); then the upper
 * triangle of B'*DB is added to Ke (column-major Kedim x Kedim). */
void orc_add_btdb_ut_only(int kd, double *Ke, const double *B, double Jac_w, const double *D, double *DB)
{
    const int ns = 6;
    for (int nx = 0; nx < kd; nx++)
        for (int mx = 0; mx < ns; mx++) {
            double acc = 0.0;
            for (int px = 0; px < ns; px++) acc += D[ns * mx + px] * B[px * kd + nx];
            DB[mx * kd + nx] = Jac_w * acc;
        }
    for (int nx = 0; nx < kd; nx++)
        for (int mx = 0; mx <= nx; mx++) {
            double acc = 0.0;
            for (int px = 0; px < ns; px++) acc += B[px * kd + mx] * DB[px * kd + nx];
            Ke[mx + kd * nx] += acc;
        }
}

/* complete_lt! -- src/MatrixUtilityModule.jl:164-173. */
void orc_complete_lt(int n, double *Ke)
{
    for (int nx = 0; nx < n; nx++)
        for (int mx = nx + 1; mx < n; mx++) Ke[mx + n * nx] = Ke[nx + n * mx];
}

/* ------------------------------------------------------------------------- */
/* Element matrices (the bodies of the bilform_* element loops)              */
/* ------------------------------------------------------------------------- */

typedef struct {
    int etype, npe, npts;
    double N[64 * 10], dN[64 * 30], w[64];
} orc_idata;

static int orc_idata_make(orc_idata *d, int etype, int rule, int order)
{
    d->etype = etype;
    d->npe = (etype == ORC_H8) ? 8 : 10;
    d->npts = orc_integration_data(etype, rule, order, d->N, d->dN, d->w);
    return d->npts;
}

/* One element matrix (column-major elmdim x elmdim) for the given form.
 *   DIFFUSION_ISO  _bilform_diffusion_iso   src/FEMMBaseModule.jl:1508-1535
 *   DIFFUSION_MAT  _bilform_diffusion_general src/FEMMBaseModule.jl:1476-1506
 *   LIN_ELASTIC    bilform_lin_elastic      src/FEMMBaseModule.jl:1774-1813
 *   DOT            bilform_dot (m = 3)      src/FEMMBaseModule.jl:1335-1366
 * coeff: ISO -> c (1 value); MAT -> kappa 3x3 row-major; LIN_ELASTIC -> D 6x6
 * row-major; DOT -> c ndn x ndn row-major.  The material frame is the identity
 * CSys(3) (src/FEMMBaseModule.jl:82-84), so RmTJ = J. */
static void orc_element_matrix(const orc_idata *d, int form, int ndn, const double *X, const double *coeff,
                               double *Ke)
{
    const int npe = d->npe;
    const int kd = ndn * npe;
    double loc[3], J[9], gradN[30], B[6 * 30], DB[6 * 30];
    static const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    memset(Ke, 0, sizeof(double) * kd * kd);
    for (int j = 0; j < d->npts; j++) {
        const double *Nj = d->N + j * npe;
        const double *dNj = d->dN + j * npe * 3;
        orc_locjac(npe, X, Nj, dNj, loc, J);
        double Jac = orc_jacobian_det(J);
        if (form == ORC_FORM_DIFFUSION_ISO) {
            orc_gradN(npe, dNj, J, gradN);
            orc_add_mggt_ut_only(npe, Ke, gradN, (coeff[0] * Jac * d->w[j]));
        } else if (form == ORC_FORM_DIFFUSION_MAT) {
            orc_gradN(npe, dNj, J, gradN);
            orc_add_gkgt_ut_only(npe, Ke, gradN, (Jac * d->w[j]), coeff);
        } else if (form == ORC_FORM_LIN_ELASTIC) {
            orc_gradN(npe, dNj, J, gradN);
            orc_blmat3d(npe, gradN, I3, B);
            orc_add_btdb_ut_only(kd, Ke, B, Jac * d->w[j], coeff, DB);
        } else { /* DOT: full (not triangular) dense fill, :1355-1360 */
            for (int k = 0; k < npe; k++)
                for (int m = 0; m < npe; m++) {
                    double factor = (Nj[k] * Nj[m] * Jac * d->w[j]);
                    for (int p = 0; p < ndn; p++)
                        for (int q = 0; q < ndn; q++)
                            Ke[(k * ndn + p) + kd * (m * ndn + q)] += factor * coeff[ndn * p + q];
                }
        }
    }
    if (form != ORC_FORM_DOT) orc_complete_lt(kd, Ke);
}

/* Exposed for element-level tests: X is npe x 3 row-major, Ke column-major. */
int orc_elmat(int etype, int rule, int order, int form, int ndn, const double *X, const double *coeff, double *Ke)
{
    orc_idata d;
    if (orc_idata_make(&d, etype, rule, order) < 0) return -1;
    orc_element_matrix(&d, form, ndn, X, coeff, Ke);
    return 0;
}

/* Full element loop of bilform_* writing the COO triples exactly as
 * assemble!(::SysmatAssemblerSparse) serializes them
 * (src/AssemblyModule.jl:250-282): element-major, then local column j, then
 * local row i.  conn: nelem x npe row-major, 1-based node numbers; xyz: nnodes
 * x 3 row-major; dofnums: nnodes x ndn row-major, 1-based (gatherdofnums!,
 * src/FieldModule.jl:304-314, node-major within the element).  I/J are 1-based.
 * Returns number of triples, or -1 (bad rule), -2 (DOF < 1), -3 (DOF > n). */
int64_t orc_bilform_coo(int etype, int rule, int order, int form, int64_t nelem, const int64_t *conn,
                        const double *xyz, int ndn, const int64_t *dofnums, int64_t nalldofs,
                        const double *coeff, int64_t *I, int64_t *J, double *V)
{
    orc_idata d;
    if (orc_idata_make(&d, etype, rule, order) < 0) return -1;
    const int npe = d.npe, kd = ndn * npe;
    double X[30], Ke[900];
    int64_t dofs[30];
    int64_t p = 0;
    for (int64_t e = 0; e < nelem; e++) {
        const int64_t *c = conn + e * npe;
        for (int a = 0; a < npe; a++)
            for (int k = 0; k < 3; k++) X[3 * a + k] = xyz[3 * (c[a] - 1) + k];
        orc_element_matrix(&d, form, ndn, X, coeff, Ke);
        for (int a = 0; a < npe; a++)
            for (int k = 0; k < ndn; k++) dofs[a * ndn + k] = dofnums[(c[a] - 1) * ndn + k];
        for (int jj = 0; jj < kd; jj++) {
            int64_t dj = dofs[jj];
            if (dj < 1) return -2;
            if (dj > nalldofs) return -3;
            for (int ii = 0; ii < kd; ii++) {
                int64_t di = dofs[ii];
                if (di < 1) return -2;
                if (di > nalldofs) return -3;
                V[p] = Ke[ii + kd * jj];
                I[p] = di;
                J[p] = dj;
                p++;
            }
        }
    }
    return p;
}

/* orc_bilform_coo with the element loop spread over nthreads OpenMP threads
 * (CPU-baseline timing only).  Every element owns its kd*kd triple slots
 * [e*kd*kd, (e+1)*kd*kd), so the threads write disjoint ranges and the
 * triples are the serial function's, in the same order; the reference's loop
 * is serial, and this variant is the multi-core CPU figure beside it.
 * Returns the number of triples or the serial function's error codes. */
int64_t orc_bilform_coo_threads(int etype, int rule, int order, int form, int64_t nelem, const int64_t *conn,
                                const double *xyz, int ndn, const int64_t *dofnums, int64_t nalldofs,
                                const double *coeff, int64_t *I, int64_t *J, double *V, int nthreads)
{
    orc_idata d;
    if (orc_idata_make(&d, etype, rule, order) < 0) return -1;
    const int npe = d.npe, kd = ndn * npe;
    int err = 0;
#pragma omp parallel for schedule(static) num_threads(nthreads) reduction(min : err)
    for (int64_t e = 0; e < nelem; e++) {
        double X[30], Ke[900];
        int64_t dofs[30];
        const int64_t *c = conn + e * npe;
        for (int a = 0; a < npe; a++)
            for (int k = 0; k < 3; k++) X[3 * a + k] = xyz[3 * (c[a] - 1) + k];
        orc_element_matrix(&d, form, ndn, X, coeff, Ke);
        for (int a = 0; a < npe; a++)
            for (int k = 0; k < ndn; k++) dofs[a * ndn + k] = dofnums[(c[a] - 1) * ndn + k];
        int64_t p = e * (int64_t)kd * kd;
        for (int jj = 0; jj < kd; jj++)
            for (int ii = 0; ii < kd; ii++) {
                const int64_t di = dofs[ii], dj = dofs[jj];
                if (di < 1 || dj < 1) err = err < -2 ? err : -2;
                else if (di > nalldofs || dj > nalldofs) err = err < -3 ? err : -3;
                V[p] = Ke[ii + kd * jj];
                I[p] = di;
                J[p] = dj;
                p++;
            }
    }
    return err < 0 ? err : nelem * (int64_t)kd * kd;
}

/* bilform_dot (src/FEMMBaseModule.jl:1335-1366) into a
 * SysmatAssemblerSparseHRZLumpingSymm (src/AssemblyModule.jl:943-1141): for
 * every element mass matrix, em2 = sum(sum(mat, dims = 1)) (column sums, then
 * their sum), dem2 = sum of the diagonal, ffactor = em2 / dem2, and the
 * triples (dj, dj, mat[j,j] * ffactor) for j in element order; makematrix!
 * is sparse() over them.  Writes npe*ndn triples per element into I/J/V
 * (1-based) and returns their count, or -1 (bad rule), -2 (DOF < 1), -3 (DOF
 * > size).  Julia's sum may reassociate (@simd), so em2's last bits are
 * "parity unpinned". */
int64_t orc_hrz_lumped_coo(int etype, int rule, int order, int64_t nelem, const int64_t *conn, const double *xyz,
                           int ndn, const int64_t *dofnums, int64_t nalldofs, const double *coeff, int64_t *I,
                           int64_t *J, double *V)
{
    orc_idata d;
    if (orc_idata_make(&d, etype, rule, order) < 0) return -1;
    const int npe = d.npe, kd = ndn * npe;
    double X[30], Ke[900];
    int64_t p = 0;
    for (int64_t e = 0; e < nelem; e++) {
        const int64_t *c = conn + e * npe;
        for (int a = 0; a < npe; a++)
            for (int k = 0; k < 3; k++) X[3 * a + k] = xyz[3 * (c[a] - 1) + k];
        orc_element_matrix(&d, ORC_FORM_DOT, ndn, X, coeff, Ke);
        double em2 = 0.0;
        for (int jj = 0; jj < kd; jj++) {
            double cs = 0.0;
            for (int ii = 0; ii < kd; ii++) cs += Ke[ii + kd * jj];
            em2 += cs;
        }
        double dem2 = 0.0;
        for (int ii = 0; ii < kd; ii++) dem2 += Ke[ii + kd * ii];
        const double ffactor = em2 / dem2;
        for (int a = 0; a < npe; a++)
            for (int k = 0; k < ndn; k++) {
                const int jj = a * ndn + k;
                const int64_t dj = dofnums[(c[a] - 1) * ndn + k];
                if (dj < 1) return -2;
                if (dj > nalldofs) return -3;
                V[p] = Ke[jj + kd * jj] * ffactor;
                I[p] = dj;
                J[p] = dj;
                p++;
            }
    }
    return p;
}

/* linform_dot with a constant force on a 3-manifold (m = 3) and a
 * SysvecAssembler (src/FEMMBaseModule.jl:1207-1241; assembler
 * src/AssemblyModule.jl:853-917).  Per element and quadrature point:
 * Factor = Jac * w (Jacobianmdim(m = 3) = det J for solids,
 * src/IntegDomainModule.jl:567-575), NkxF = N_k * Factor, elvec[(k, mx)] +=
 * NkxF * force[mx]; then F[dofnums[i]] += elvec[i] element by element, on a
 * buffer zeroed by startassembly!.  conn 1-based (nelem x npe), xyz nnodes x
 * 3, dofnums nnodes x ndn 1-based; F has nalldofs entries (0-based storage).
 * Returns 0, -1 (bad rule), -2 (DOF < 1), -3 (DOF > size). */
int orc_linform_dot(int etype, int rule, int order, int64_t nelem, const int64_t *conn, const double *xyz, int ndn,
                    const int64_t *dofnums, int64_t nalldofs, const double *force, double *F)
{
    orc_idata d;
    if (orc_idata_make(&d, etype, rule, order) < 0) return -1;
    const int npe = d.npe;
    double X[30], loc[3], J[9], elvec[30];
    for (int64_t i = 0; i < nalldofs; i++) F[i] = 0.0;
    for (int64_t e = 0; e < nelem; e++) {
        const int64_t *c = conn + e * npe;
        for (int a = 0; a < npe; a++)
            for (int k = 0; k < 3; k++) X[3 * a + k] = xyz[3 * (c[a] - 1) + k];
        for (int i = 0; i < npe * ndn; i++) elvec[i] = 0.0;
        for (int j = 0; j < d.npts; j++) {
            const double *Nj = d.N + j * npe;
            orc_locjac(npe, X, Nj, d.dN + j * npe * 3, loc, J);
            const double Jac = orc_jacobian_det(J);
            const double Factor = Jac * d.w[j];
            int rx = 0;
            for (int kx = 0; kx < npe; kx++) {
                const double NkxF = Nj[kx] * Factor;
                for (int mx = 0; mx < ndn; mx++) {
                    elvec[rx] = elvec[rx] + NkxF * force[mx];
                    rx++;
                }
            }
        }
        for (int a = 0; a < npe; a++)
            for (int k = 0; k < ndn; k++) {
                const int64_t gi = dofnums[(c[a] - 1) * ndn + k];
                if (gi < 1) return -2;
                if (gi > nalldofs) return -3;
                F[gi - 1] += elvec[a * ndn + k];
            }
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* sparse(I, J, V, m, n) -- Julia stdlib SparseArrays (called from            */
/* makematrix!, src/AssemblyModule.jl:319-325).  Not in the reference tree;  */
/* restated from the stdlib's published algorithm: a stable counting sort of */
/* the triples into CSR by row, duplicate columns within a row combined with */
/* + as a left fold in input order, explicit zeros retained, then a          */
/* transpose to CSC, which leaves the rows of every column sorted.           */
/* ------------------------------------------------------------------------- */

typedef struct {
    int64_t m, n, nnz;
    int64_t *colptr; /* 1-based, n+1 */
    int64_t *rowval; /* 1-based */
    double *nzval;
} orc_csc;

orc_csc *orc_sparse(int64_t ntrip, const int64_t *I, const int64_t *J, const double *V, int64_t m, int64_t n)
{
    for (int64_t k = 0; k < ntrip; k++)
        if (I[k] < 1 || I[k] > m || J[k] < 1 || J[k] > n) return NULL;
    int64_t *rowptr = calloc((size_t)m + 2, sizeof(int64_t));
    int64_t *ccol = malloc(sizeof(int64_t) * (size_t)(ntrip ? ntrip : 1));
    double *cval = malloc(sizeof(double) * (size_t)(ntrip ? ntrip : 1));
    /* counting sort by row (stable) */
    for (int64_t k = 0; k < ntrip; k++) rowptr[I[k] + 1]++;
    rowptr[0] = 0; /* rowptr[i] = start of row i (1-based rows, index i) */
    rowptr[1] = 0;
    for (int64_t i = 2; i <= m + 1; i++) rowptr[i] += rowptr[i - 1];
    /* rowptr[i] now = start of row i-1 ... shift: use next[] as insertion cursors */
    int64_t *next = malloc(sizeof(int64_t) * ((size_t)m + 2));
    for (int64_t i = 1; i <= m + 1; i++) next[i] = rowptr[i];
    for (int64_t k = 0; k < ntrip; k++) {
        int64_t pos = next[I[k]]++;
        ccol[pos] = J[k];
        cval[pos] = V[k];
    }
    /* rows occupy [rowptr[i], rowptr[i+1]) for i = 1..m */
    int64_t *lasttouch = malloc(sizeof(int64_t) * ((size_t)n + 1));
    for (int64_t j = 0; j <= n; j++) lasttouch[j] = -1;
    int64_t *colcnt = calloc((size_t)n + 2, sizeof(int64_t));
    int64_t *keep_row_end = malloc(sizeof(int64_t) * ((size_t)m + 2));
    int64_t w = 0;
    for (int64_t i = 1; i <= m; i++) {
        int64_t rstart = w;
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++) {
            int64_t j = ccol[k];
            int64_t kl = lasttouch[j];
            if (kl >= rstart) {
                cval[kl] = cval[kl] + cval[k]; /* left fold in input order */
            } else {
                ccol[w] = j;
                cval[w] = cval[k];
                lasttouch[j] = w;
                colcnt[j + 1]++;
                w++;
            }
        }
        keep_row_end[i] = w;
    }
    orc_csc *S = malloc(sizeof(orc_csc));
    S->m = m;
    S->n = n;
    S->nnz = w;
    S->colptr = malloc(sizeof(int64_t) * ((size_t)n + 1));
    S->rowval = malloc(sizeof(int64_t) * (size_t)(w ? w : 1));
    S->nzval = malloc(sizeof(double) * (size_t)(w ? w : 1));
    /* transpose CSR -> CSC */
    int64_t *cnext = malloc(sizeof(int64_t) * ((size_t)n + 2));
    cnext[1] = 0;
    for (int64_t j = 1; j <= n; j++) cnext[j + 1] = cnext[j] + colcnt[j + 1];
    for (int64_t j = 1; j <= n + 1; j++) S->colptr[j - 1] = cnext[j] + 1;
    int64_t rs = 0;
    for (int64_t i = 1; i <= m; i++) {
        for (int64_t k = rs; k < keep_row_end[i]; k++) {
            int64_t j = ccol[k];
            int64_t pos = cnext[j]++;
            S->rowval[pos] = i;
            S->nzval[pos] = cval[k];
        }
        rs = keep_row_end[i];
    }
    free(rowptr);
    free(ccol);
    free(cval);
    free(next);
    free(lasttouch);
    free(colcnt);
    free(keep_row_end);
    free(cnext);
    return S;
}

int64_t orc_csc_nnz(const orc_csc *S) { return S->nnz; }

void orc_csc_copy(const orc_csc *S, int64_t *colptr, int64_t *rowval, double *nzval)
{
    memcpy(colptr, S->colptr, sizeof(int64_t) * (size_t)(S->n + 1));
    memcpy(rowval, S->rowval, sizeof(int64_t) * (size_t)S->nnz);
    memcpy(nzval, S->nzval, sizeof(double) * (size_t)S->nnz);
}

void orc_csc_free(orc_csc *S)
{
    if (!S) return;
    free(S->colptr);
    free(S->rowval);
    free(S->nzval);
    free(S);
}

/* makematrix!(::SysmatAssemblerSparseSymm) after sparse() of the lower-
 * triangle triples (src/AssemblyModule.jl:567-583): S + transpose(S), where
 * SparseArrays' + stores only nonzero results, then every stored diagonal
 * entry halved (S[j, j] *= 0.5 leaves an absent entry absent).  Returns a new
 * CSC (1-based). */
orc_csc *orc_symm_from_lower(const orc_csc *S)
{
    const int64_t n = S->n;
    /* transpose of S by a counting pass */
    int64_t *tp = calloc((size_t)n + 2, sizeof(int64_t));
    int64_t *tr = malloc(sizeof(int64_t) * (size_t)(S->nnz ? S->nnz : 1));
    double *tv = malloc(sizeof(double) * (size_t)(S->nnz ? S->nnz : 1));
    for (int64_t k = 0; k < S->nnz; k++) tp[S->rowval[k]]++; /* 1-based row r counts in slot r */
    for (int64_t c = 0; c < n; c++) tp[c + 1] += tp[c];
    int64_t *cur = malloc(sizeof(int64_t) * (size_t)(n + 1));
    for (int64_t c = 0; c <= n; c++) cur[c] = tp[c];
    for (int64_t c = 0; c < n; c++)
        for (int64_t k = S->colptr[c] - 1; k < S->colptr[c + 1] - 1; k++) {
            const int64_t r = S->rowval[k] - 1; /* row of S = column of the transpose */
            tr[cur[r]] = c + 1;
            tv[cur[r]] = S->nzval[k];
            cur[r]++;
        }
    orc_csc *C = malloc(sizeof(orc_csc));
    C->m = S->m;
    C->n = n;
    C->colptr = malloc(sizeof(int64_t) * (size_t)(n + 1));
    C->rowval = malloc(sizeof(int64_t) * (size_t)(2 * S->nnz + 1));
    C->nzval = malloc(sizeof(double) * (size_t)(2 * S->nnz + 1));
    int64_t q = 0;
    C->colptr[0] = 1;
    for (int64_t c = 0; c < n; c++) {
        int64_t a = S->colptr[c] - 1, ae = S->colptr[c + 1] - 1, b = tp[c], be = tp[c + 1];
        while (a < ae || b < be) {
            int64_t r;
            double v;
            if (b >= be || (a < ae && S->rowval[a] < tr[b])) {
                r = S->rowval[a];
                v = S->nzval[a] + 0.0;
                a++;
            } else if (a >= ae || tr[b] < S->rowval[a]) {
                r = tr[b];
                v = 0.0 + tv[b];
                b++;
            } else {
                r = S->rowval[a];
                v = S->nzval[a] + tv[b];
                a++;
                b++;
            }
            if (v != 0.0) {
                C->rowval[q] = r;
                C->nzval[q] = (r == c + 1) ? v * 0.5 : v;
                q++;
            }
        }
        C->colptr[c + 1] = q + 1;
    }
    C->nnz = q;
    free(tp);
    free(tr);
    free(tv);
    free(cur);
    return C;
}

/* Element loop of bilform_* into a SysmatAssemblerSparseSymm: assemble!
 * keeps the lower triangle in element-local order, for j, for i = j..kd
 * (src/AssemblyModule.jl:494-523).  Same arguments and returns as
 * orc_bilform_coo. */
int64_t orc_bilform_coo_symm(int etype, int rule, int order, int form, int64_t nelem, const int64_t *conn,
                             const double *xyz, int ndn, const int64_t *dofnums, int64_t nalldofs,
                             const double *coeff, int64_t *I, int64_t *J, double *V)
{
    orc_idata d;
    if (orc_idata_make(&d, etype, rule, order) < 0) return -1;
    const int npe = d.npe, kd = ndn * npe;
    double X[30], Ke[900];
    int64_t dofs[30];
    int64_t p = 0;
    for (int64_t e = 0; e < nelem; e++) {
        const int64_t *c = conn + e * npe;
        for (int a = 0; a < npe; a++)
            for (int k = 0; k < 3; k++) X[3 * a + k] = xyz[3 * (c[a] - 1) + k];
        orc_element_matrix(&d, form, ndn, X, coeff, Ke);
        for (int a = 0; a < npe; a++)
            for (int k = 0; k < ndn; k++) dofs[a * ndn + k] = dofnums[(c[a] - 1) * ndn + k];
        for (int jj = 0; jj < kd; jj++) {
            const int64_t dj = dofs[jj];
            if (dj < 1) return -2;
            if (dj > nalldofs) return -3;
            for (int ii = jj; ii < kd; ii++) {
                const int64_t di = dofs[ii];
                if (di < 1) return -2;
                if (di > nalldofs) return -3;
                V[p] = Ke[ii + kd * jj];
                I[p] = di;
                J[p] = dj;
                p++;
            }
        }
    }
    return p;
}

/* ------------------------------------------------------------------------- */
/* Meshes                                                                    */
/* ------------------------------------------------------------------------- */

/* H8blockx -- src/MeshHexahedronModule.jl:59-112.  Nodes x fastest (:76-86);
 * cells with k (z) innermost (:88-105); local order f, f+1, f+nx+2, f+nx+1,
 * then the same one layer up.  conn 1-based nelem x 8, xyz nnodes x 3. */
void orc_h8blockx(const double *xs, int64_t nxs, const double *ys, int64_t nys, const double *zs, int64_t nzs,
                  int64_t *conn, double *xyz)
{
    int64_t nL = nxs - 1, nW = nys - 1, nH = nzs - 1;
    int64_t f = 0;
    for (int64_t k = 0; k < nH + 1; k++)
        for (int64_t j = 0; j < nW + 1; j++)
            for (int64_t i = 0; i < nL + 1; i++) {
                xyz[3 * f + 0] = xs[i];
                xyz[3 * f + 1] = ys[j];
                xyz[3 * f + 2] = zs[k];
                f++;
            }
    int64_t gc = 0;
    for (int64_t i = 1; i <= nL; i++)
        for (int64_t j = 1; j <= nW; j++)
            for (int64_t k = 1; k <= nH; k++) {
                int64_t ff = (k - 1) * ((nL + 1) * (nW + 1)) + (j - 1) * (nL + 1) + i;
                int64_t *c = conn + 8 * gc;
                c[0] = ff;
                c[1] = ff + 1;
                c[2] = ff + (nL + 1) + 1;
                c[3] = ff + (nL + 1);
                ff = ff + (nL + 1) * (nW + 1);
                c[4] = ff;
                c[5] = ff + 1;
                c[6] = ff + (nL + 1) + 1;
                c[7] = ff + (nL + 1);
                gc++;
            }
}

/* T4blockx(orientation = :a) -- src/MeshTetrahedronModule.jl:81-151: six tets
 * per cell from the table t4ia = t4ib (:94-96); both parities use the same
 * table for :a.  conn 1-based nelem x 4. */
void orc_t4blockx_a(const double *xs, int64_t nxs, const double *ys, int64_t nys, const double *zs, int64_t nzs,
                    int64_t *conn, double *xyz)
{
    static const int t4[6][4] = {{1, 8, 5, 6}, {3, 4, 2, 7}, {7, 2, 6, 8}, {4, 7, 8, 2}, {2, 1, 6, 8}, {4, 8, 1, 2}};
    int64_t nL = nxs - 1, nW = nys - 1, nH = nzs - 1;
    int64_t f = 0;
    for (int64_t k = 0; k < nH + 1; k++)
        for (int64_t j = 0; j < nW + 1; j++)
            for (int64_t i = 0; i < nL + 1; i++) {
                xyz[3 * f + 0] = xs[i];
                xyz[3 * f + 1] = ys[j];
                xyz[3 * f + 2] = zs[k];
                f++;
            }
    int64_t gc = 0;
    for (int64_t i = 1; i <= nL; i++)
        for (int64_t j = 1; j <= nW; j++)
            for (int64_t k = 1; k <= nH; k++) {
                int64_t ff = (k - 1) * ((nL + 1) * (nW + 1)) + (j - 1) * (nL + 1) + i;
                int64_t nn[8] = {ff, ff + 1, ff + (nL + 1) + 1, ff + (nL + 1), 0, 0, 0, 0};
                for (int t = 0; t < 4; t++) nn[4 + t] = nn[t] + (nL + 1) * (nW + 1);
                for (int r = 0; r < 6; r++) {
                    for (int c1 = 0; c1 < 4; c1++) conn[4 * gc + c1] = nn[t4[r][c1] - 1];
                    gc++;
                }
            }
}

/* T4toT10 -- src/MeshTetrahedronModule.jl:158-212 with the hyper-face container
 * of src/MeshUtilModule.jl:41-80: local edges (1,2),(2,3),(3,1),(4,1),(4,2),(4,3)
 * (:160); a new node number is issued on the first encounter of an edge, in
 * element order then local-edge order; the mid-edge node is the mean of the
 * two end points (:177-184).  The container is restated as an open-addressing
 * table keyed by the sorted end points.  xyz10 must hold the output node
 * count (query with xyz10 == NULL); returns the total node count. */
typedef struct {
    int64_t lo, hi, n;
} orc_edge;

static uint64_t orc_hash2(int64_t a, int64_t b)
{
    uint64_t h = ((uint64_t)a * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)b + 0x632BE59BD9B4E019ull);
    h ^= h >> 29;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 32;
    return h;
}

int64_t orc_t4tot10(int64_t nnodes4, const double *xyz4, int64_t nelem, const int64_t *conn4, int64_t *conn10,
                    double *xyz10)
{
    static const int ec[6][2] = {{1, 2}, {2, 3}, {3, 1}, {4, 1}, {4, 2}, {4, 3}};
    size_t cap = 1;
    while (cap < (size_t)(nelem * 6) * 2 + 16) cap <<= 1;
    orc_edge *tab = malloc(sizeof(orc_edge) * cap);
    for (size_t i = 0; i < cap; i++) tab[i].lo = 0;
    int64_t newn = nnodes4 + 1;
    for (int64_t e = 0; e < nelem; e++)
        for (int J = 0; J < 6; J++) {
            int64_t a = conn4[4 * e + ec[J][0] - 1], b = conn4[4 * e + ec[J][1] - 1];
            int64_t lo = a < b ? a : b, hi = a < b ? b : a;
            size_t s = orc_hash2(lo, hi) & (cap - 1);
            while (tab[s].lo != 0 && !(tab[s].lo == lo && tab[s].hi == hi)) s = (s + 1) & (cap - 1);
            if (tab[s].lo == 0) {
                tab[s].lo = lo;
                tab[s].hi = hi;
                tab[s].n = newn++;
            }
            conn10[10 * e + 4 + J] = tab[s].n;
        }
    int64_t total = newn - 1;
    if (xyz10) {
        memcpy(xyz10, xyz4, sizeof(double) * 3 * (size_t)nnodes4);
        for (size_t s = 0; s < cap; s++)
            if (tab[s].lo != 0) {
                int64_t lo = tab[s].lo, hi = tab[s].hi, n = tab[s].n;
                for (int k = 0; k < 3; k++)
                    xyz10[3 * (n - 1) + k] = (xyz10[3 * (hi - 1) + k] + xyz10[3 * (lo - 1) + k]) / 2;
            }
    }
    for (int64_t e = 0; e < nelem; e++)
        for (int a = 0; a < 4; a++) conn10[10 * e + a] = conn4[4 * e + a];
    free(tab);
    return total;
}

/* Seeded perturbation of block meshes (defined by this build, not by the
 * reference): node n (0-based) component c moves by u * amp, u in [-1, 1)
 * drawn from SplitMix64 at counter 3n + c + 1 of the stream seeded with
 * `seed`; only nodes strictly inside the box [lo, hi] (tested on the
 * unperturbed coordinates) move. */
static uint64_t orc_splitmix(uint64_t seed, uint64_t ctr)
{
    uint64_t z = seed + ctr * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void orc_perturb(int64_t nnodes, double *xyz, const double *lo, const double *hi, double amp, uint64_t seed)
{
    for (int64_t n = 0; n < nnodes; n++) {
        double *x = xyz + 3 * n;
        int inside = 1;
        for (int c = 0; c < 3; c++)
            if (!(x[c] > lo[c] && x[c] < hi[c])) inside = 0;
        if (!inside) continue;
        for (int c = 0; c < 3; c++) {
            uint64_t z = orc_splitmix(seed, 3 * (uint64_t)n + (uint64_t)c + 1);
            double u = (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
            x[c] = x[c] + u * amp;
        }
    }
}

/* Minimum det J over all elements and quadrature points (the perturbation
 * generator must keep every element valid; the reference never checks). */
double orc_min_detj(int etype, int rule, int order, int64_t nelem, const int64_t *conn, const double *xyz)
{
    orc_idata d;
    if (orc_idata_make(&d, etype, rule, order) < 0) return NAN;
    double X[30], loc[3], J[9], mn = INFINITY;
    for (int64_t e = 0; e < nelem; e++) {
        for (int a = 0; a < d.npe; a++)
            for (int k = 0; k < 3; k++) X[3 * a + k] = xyz[3 * (conn[e * d.npe + a] - 1) + k];
        for (int q = 0; q < d.npts; q++) {
            orc_locjac(d.npe, X, d.N + q * d.npe, d.dN + q * d.npe * 3, loc, J);
            double dj = orc_jacobian_det(J);
            if (dj < mn) mn = dj;
        }
    }
    return mn;
}
