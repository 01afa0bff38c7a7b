"""Pins for the oracle's element matrices (the bilform_* element loops).

Closed forms (exactly integrated unit-cube Laplacian, rigid-body null space,
mass = rho * V), the reference's patch tests (test/test_forms.jl) and an
independent matrix-algebra formulation (B^T D B by numpy products).
"""
import numpy as np
import pytest

import oracle as O

CUBE = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                 [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)


def distorted_h8(seed, amp=0.15):
    rng = np.random.default_rng(seed)
    return CUBE + rng.uniform(-amp, amp, CUBE.shape)


def distorted_t10(seed, amp=0.05):
    rng = np.random.default_rng(seed)
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) + rng.uniform(-0.1, 0.1, (4, 3))
    e = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
    m = np.array([(v[a] + v[b]) / 2 for a, b in e]) + rng.uniform(-amp, amp, (6, 3))
    return np.vstack([v, m])


def test_unit_cube_laplacian_closed_form():
    # 2x2x2 Gauss integrates the trilinear Laplacian exactly: diagonal h/3,
    # edge neighbours 0, face and body diagonals -h/12 (h = edge length)
    for h in (1.0, 0.1):
        K = O.elmat(O.H8, O.DIFFUSION_ISO, 1, CUBE * h, [1.0])
        for a in range(8):
            for b in range(8):
                d = int(np.abs(CUBE[a] - CUBE[b]).sum())
                expect = {0: h / 3, 1: 0.0, 2: -h / 12, 3: -h / 12}[d]
                assert abs(K[a, b] - expect) < 1e-14 * max(1, h), (a, b, K[a, b], expect)


def test_diffusion_scaling_and_matrix_form():
    X = distorted_h8(2)
    K1 = O.elmat(O.H8, O.DIFFUSION_ISO, 1, X, [1.0])
    K3 = O.elmat(O.H8, O.DIFFUSION_ISO, 1, X, [3.0])
    Km = O.elmat(O.H8, O.DIFFUSION_MAT, 1, X, np.eye(3).ravel())
    assert np.allclose(K3, 3 * K1, rtol=1e-14)
    assert np.allclose(Km, K1, rtol=1e-13, atol=1e-15)
    assert np.array_equal(K1, K1.T)  # complete_lt! makes it exactly symmetric
    # constant field -> zero energy, linear field -> |grad|^2 V
    assert np.abs(K1 @ np.ones(8)).max() < 1e-14
    # anisotropic kappa against the numpy formulation gradN kappa gradN^T J w
    rng = np.random.default_rng(5)
    A = rng.normal(size=(3, 3))
    kap = A @ A.T + np.eye(3)
    Kk = O.elmat(O.H8, O.DIFFUSION_MAT, 1, X, kap.ravel())
    ref = np.zeros((8, 8))
    pc, w = O.gauss_rule3(2)
    for q in range(8):
        dN = O.bfundpar(O.H8, pc[q])
        J = X.T @ dN
        g = dN @ np.linalg.inv(J)
        ref += g @ kap @ g.T * np.linalg.det(J) * w[q]
    assert np.allclose(Kk, ref, rtol=1e-12, atol=1e-14)


def iso_D(E, nu):
    return O.iso_D(E, nu)


def rigid_modes(X):
    R = []
    for k in range(3):
        u = np.zeros_like(X)
        u[:, k] = 1
        R.append(u.ravel())
    for (i, j) in ((0, 1), (0, 2), (1, 2)):
        u = np.zeros_like(X)
        u[:, i] = -X[:, j]
        u[:, j] = X[:, i]
        R.append(u.ravel())
    return np.array(R).T


@pytest.mark.parametrize("etype,X", [(O.H8, distorted_h8(11)), (O.T10, distorted_t10(12))])
def test_elasticity_rigid_null_space_and_numpy_btdb(etype, X):
    D = iso_D(200e3, 0.3)
    K = O.elmat(etype, O.LIN_ELASTIC, 3, X, D.ravel())
    assert np.array_equal(K, K.T)
    R = rigid_modes(X)
    assert np.abs(K @ R).max() < 1e-9 * np.abs(K).max()
    # independent formulation: B^T D B J w with B built by numpy
    pc, w = O.gauss_rule3(2) if etype == O.H8 else O.tet_rule(4)
    npe = X.shape[0]
    ref = np.zeros((3 * npe, 3 * npe))
    for q in range(len(w)):
        dN = O.bfundpar(etype, pc[q])
        J = X.T @ dN
        g = dN @ np.linalg.inv(J)
        B = np.zeros((6, 3 * npe))
        B[0, 0::3], B[1, 1::3], B[2, 2::3] = g[:, 0], g[:, 1], g[:, 2]
        B[3, 0::3], B[3, 1::3] = g[:, 1], g[:, 0]
        B[4, 0::3], B[4, 2::3] = g[:, 2], g[:, 0]
        B[5, 1::3], B[5, 2::3] = g[:, 2], g[:, 1]
        ref += B.T @ D @ B * np.linalg.det(J) * w[q]
    assert np.allclose(K, ref, rtol=1e-11, atol=1e-9)


def test_elasticity_anisotropic_D_against_numpy():
    rng = np.random.default_rng(9)
    A = rng.normal(size=(6, 6))
    D = A @ A.T + 6 * np.eye(6)
    X = distorted_h8(4)
    K = O.elmat(O.H8, O.LIN_ELASTIC, 3, X, D.ravel())
    pc, w = O.gauss_rule3(2)
    ref = np.zeros((24, 24))
    for q in range(8):
        dN = O.bfundpar(O.H8, pc[q])
        J = X.T @ dN
        B = O.blmat3d(dN @ np.linalg.inv(J))
        ref += B.T @ D @ B * np.linalg.det(J) * w[q]
    assert np.allclose(K, ref, rtol=1e-11, atol=1e-9)


@pytest.mark.parametrize("etype,X", [(O.H8, distorted_h8(21)), (O.T10, distorted_t10(22))])
@pytest.mark.parametrize("ndn", [1, 3])
def test_mass_sums_to_rho_volume(etype, X, ndn):
    rho = 2.7e-9
    c = rho * np.eye(ndn)
    M = O.elmat(etype, O.DOT, ndn, X, c.ravel())
    assert np.array_equal(M, M.T)
    V = O.elmat(etype, O.DOT, 1, X, [1.0]).sum()
    pc, w = O.gauss_rule3(2) if etype == O.H8 else O.tet_rule(4)
    vol = sum(np.linalg.det(X.T @ O.bfundpar(etype, pc[q])) * w[q] for q in range(len(w)))
    assert abs(V - vol) < 1e-14 * abs(vol)
    for k in range(ndn):
        assert abs(M[k::ndn, k::ndn].sum() - rho * vol) < 1e-12 * rho * vol
    if ndn == 3:
        assert M[0::3, 1::3].sum() == 0.0


def test_patch_energies_like_reference_forms():
    # test/test_forms.jl:54-78 (diffusion, linear field) and :331-367
    # (elasticity, linear displacement), on a perturbed block
    conn, xyz = O.block_mesh(O.H8, (3, 4, 2), L=1.0, amp_frac=0.1, seed=3)
    V = 1.0
    b, c, d = 0.3, 0.4, -0.5
    psi = -0.1 + b * xyz[:, 0] + c * xyz[:, 1] + d * xyz[:, 2]
    cp, rv, nz = O.assemble_csc(O.H8, O.DIFFUSION_ISO, conn, xyz, 1, [1.0])
    import scipy.sparse as sp
    n = xyz.shape[0]
    K = sp.csc_matrix((nz, rv - 1, cp - 1), shape=(n, n))
    assert abs(psi @ (K @ psi) - (b * b + c * c + d * d) * V) / V <= 1e-5
    assert abs(np.ones(n) @ (K @ np.ones(n))) <= 1e-12
    mu = 0.13377
    D = np.diag([2 * mu] * 3 + [mu] * 3)
    a_, b_, c_, d_ = (-0.33, 2 / 3, -1.67, 2 / 7)
    X, Y, Z = xyz.T
    u = np.column_stack([a_ + b_ * X + c_ * Y + d_ * Z, b_ + c_ * X + d_ * Y + a_ * Z, c_ + d_ * X + a_ * Y + b_ * Z])
    gradu = np.array([[b_, c_, d_], [c_, d_, a_], [d_, a_, b_]])
    sym = (gradu + gradu.T) / 2
    cp, rv, nz = O.assemble_csc(O.H8, O.LIN_ELASTIC, conn, xyz, 3, D.ravel())
    G = sp.csc_matrix((nz, rv - 1, cp - 1), shape=(3 * n, 3 * n))
    v = u.ravel()
    exact = 2 * mu * V * (sym ** 2).sum()
    assert abs(v @ (G @ v) - exact) / exact <= 1e-5
    # test/test_forms.jl:442,485,527: bit-exact symmetry
    assert (G - G.T).count_nonzero() == 0
