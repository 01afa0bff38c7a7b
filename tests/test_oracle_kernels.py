"""Pins for the oracle's quadrature, basis and per-point kernels.

Checks come from the reference's own tests (golden gradN), from the source
literals, and from mathematics (exactness, partition of unity, affine maps),
chosen so that a dropped term, a wrong sign or a transposed operand fails.
"""
import numpy as np
import pytest

import oracle as O


def test_gauss_rule_literals_and_order():
    # src/IntegRuleModule.jl:212-213 literal, k (zeta) fastest (:381-389)
    pc, w = O.gauss_rule3(2)
    a = 0.577350269189626
    assert pc[0].tolist() == [-a, -a, -a]
    assert pc[1].tolist() == [-a, -a, a]
    assert pc[2].tolist() == [-a, a, -a]
    assert pc[4].tolist() == [a, -a, -a]
    assert np.all(w == 1.0)
    # the literal, not 1/sqrt(3)
    assert a != 1 / np.sqrt(3)


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_gauss_exactness(order):
    pc, w = O.gauss_rule3(order)
    assert abs(w.sum() - 8.0) < 1e-12
    # exact for per-axis degree <= 2*order-1 (literals have ~15 digits)
    for px in range(2 * order):
        for py in range(2 * order):
            f = pc[:, 0] ** px * pc[:, 1] ** py * pc[:, 2] ** (2 * order - 1 - (px % 2))
            exact = np.prod([0 if p % 2 else 2 / (p + 1) for p in (px, py, 2 * order - 1 - (px % 2))])
            assert abs(np.dot(w, f) - exact) < 1e-13


def test_tet_rule_literals():
    # src/IntegRuleModule.jl:488-493: the 8-digit literals as typed
    pc, w = O.tet_rule(4)
    assert pc[0].tolist() == [0.13819660] * 3
    assert pc[1].tolist() == [0.58541020, 0.13819660, 0.13819660]
    assert pc[3].tolist() == [0.13819660, 0.13819660, 0.58541020]
    assert np.all(w == 0.041666666666666666667)
    # linear monomials integrate exactly over the unit tet (1/24 each)
    for k in range(3):
        assert abs(np.dot(w, pc[:, k]) - 1 / 24) < 1e-9
    # quadratics only to the truncation of the literals (about 1e-9)
    assert abs(np.dot(w, pc[:, 0] ** 2) - 1 / 60) < 1e-8
    pc5, w5 = O.tet_rule(5)
    assert abs(w5.sum() - 1 / 6) < 1e-15 and w5[0] < 0
    pc1, w1 = O.tet_rule(1)
    assert pc1[0].tolist() == [0.25] * 3


def test_h8_golden_gradN():
    # test/test_basics.jl:1175-1211 (mgrad3mm): gradN at (0.57, 0.57, -0.57), J = I
    ref = np.array([
        [-0.0843875, -0.0843875, -0.023112500000000005],
        [0.0843875, -0.30811249999999996, -0.0843875],
        [0.30811249999999996, 0.30811249999999996, -0.30811249999999996],
        [-0.30811249999999996, 0.0843875, -0.0843875],
        [-0.023112500000000005, -0.023112500000000005, 0.023112500000000005],
        [0.023112500000000005, -0.0843875, 0.0843875],
        [0.0843875, 0.0843875, 0.30811249999999996],
        [-0.0843875, 0.023112500000000005, 0.0843875]])
    g = O.gradN(O.bfundpar(O.H8, [0.57, 0.57, -0.57]), np.eye(3))
    assert np.linalg.norm(g - ref) <= 1e-6
    assert np.array_equal(g, ref)  # and in fact bit for bit


@pytest.mark.parametrize("etype", [O.H8, O.T10])
def test_basis_partition_delta_and_gradient_sum(etype):
    rng = np.random.default_rng(3)
    if etype == O.H8:
        nodes = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                          [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)
    else:
        v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
        e = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
        nodes = np.vstack([v, [(v[a] + v[b]) / 2 for a, b in e]])
    for x in nodes:
        N = O.bfun(etype, x)
        assert np.allclose(N, (np.abs(nodes - x).sum(1) < 1e-12).astype(float), atol=1e-15)
    for _ in range(20):
        x = rng.uniform(-0.9, 0.9, 3) if etype == O.H8 else rng.dirichlet([1, 1, 1, 1])[:3]
        assert abs(O.bfun(etype, x).sum() - 1) < 1e-14
        dN = O.bfundpar(etype, x)
        assert np.abs(dN.sum(0)).max() < 1e-13
        # derivative check by central differences
        h = 1e-6
        for k in range(3):
            d = np.zeros(3)
            d[k] = h
            fd = (O.bfun(etype, x + d) - O.bfun(etype, x - d)) / (2 * h)
            assert np.abs(fd - dN[:, k]).max() < 1e-8


def test_locjac_det_and_gradN_affine():
    rng = np.random.default_rng(7)
    A = rng.normal(size=(3, 3)) + 3 * np.eye(3)
    t = rng.normal(size=3)
    ref = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                    [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)
    X = ref @ A.T + t  # affine image: x = A xi + t
    xi = rng.uniform(-1, 1, 3)
    N, dN = O.bfun(O.H8, xi), O.bfundpar(O.H8, xi)
    loc, J = O.locjac(X, N, dN)
    assert np.allclose(J, A, rtol=1e-14, atol=1e-14)  # J[i][k] = dx_i/dxi_k
    assert np.allclose(loc, A @ xi + t, atol=1e-14)
    assert abs(O.jacobian_det(J) - np.linalg.det(A)) < 1e-12 * abs(np.linalg.det(A))
    g = O.gradN(dN, J)
    assert np.allclose(g, dN @ np.linalg.inv(J), rtol=1e-13, atol=1e-14)
    # gradient of a linear field f(x) = c.x is c
    c = rng.normal(size=3)
    assert np.allclose(g.T @ (X @ c), c, atol=1e-12)


def test_blmat_structure():
    rng = np.random.default_rng(1)
    g = rng.normal(size=(8, 3))
    B = O.blmat3d(g)
    for a in range(8):
        gx, gy, gz = g[a]
        blk = B[:, 3 * a:3 * a + 3]
        expect = np.array([[gx, 0, 0], [0, gy, 0], [0, 0, gz], [gy, gx, 0], [gz, 0, gx], [0, gz, gy]])
        assert np.array_equal(blk, expect)
