"""Pins for the oracle's linform_dot + SysvecAssembler restatement
(src/FEMMBaseModule.jl:1207-1241, src/AssemblyModule.jl:853-917) from closed
forms: exact nodal shares of a constant body load on affine elements, the
total load f * volume on distorted meshes, and the reference's DOF errors."""
import numpy as np
import pytest

import oracle as O


def test_h8_affine_cube_nodal_shares_are_exact():
    # 2x2x2 Gauss integrates the trilinear N_a exactly: every element gives
    # each of its nodes V_e / 8 of the load, so a node's entry is
    # (incident elements) * V_e / 8 * f.
    n = 3
    conn, xyz = O.block_mesh(O.H8, (n, n, n), L=1.0)
    f = np.array([1.0, -2.0, 0.5])
    F = O.linform_dot(O.H8, conn, xyz, 3, f).reshape(-1, 3)
    Ve = 1.0 / n ** 3
    count = np.bincount(conn.ravel() - 1, minlength=xyz.shape[0])
    np.testing.assert_allclose(F, (count * Ve / 8)[:, None] * f[None, :], rtol=1e-14, atol=1e-17)


def test_total_load_is_force_times_volume_on_distorted_meshes():
    # sum_a N_a = 1, and det J is integrated exactly by 2x2x2 Gauss on
    # trilinear hexahedra (degree <= 2 per variable) and by TetRule(4) on
    # straight-sided tetrahedra (constant); boundary nodes are not moved, so
    # the block's volume stays 1.
    f = np.array([3.0, 0.25, -1.5])
    conn, xyz = O.block_mesh(O.H8, (4, 3, 5), L=1.0, amp_frac=0.1, seed=7)
    F = O.linform_dot(O.H8, conn, xyz, 3, f).reshape(-1, 3)
    np.testing.assert_allclose(F.sum(axis=0), f, rtol=1e-12)
    g = O.linearspace(0.0, 1.0, 4)
    conn4, xyz4 = O.t4blockx_a(g, g, g)
    xyz4 = O.perturb(xyz4, [0.0, 0.0, 0.0], [1.0, 1.0, 1.0], 0.1 / 3, 7)
    conn10, xyz10 = O.t4tot10(conn4, xyz4)  # mid-edge nodes at the midpoints
    F = O.linform_dot(O.T10, conn10, xyz10, 3, f).reshape(-1, 3)
    np.testing.assert_allclose(F.sum(axis=0), f, rtol=1e-12)


def test_curved_t10_total_load_is_the_rule_volume():
    # with independently moved mid-edge nodes det J is cubic and TetRule(4)
    # is no longer exact: the total is f times the rule's volume, which the
    # vector must reproduce (sum over elements of sum_q det J w)
    conn, xyz = O.block_mesh(O.T10, (3, 2, 3), L=1.0, amp_frac=0.1, seed=7)
    F = O.linform_dot(O.T10, conn, xyz, 1, [1.0])
    pc, w = O.tet_rule(4)
    V = 0.0
    for e in conn:
        X = xyz[e - 1]
        for q in range(4):
            dN = O.bfundpar(O.T10, pc[q])
            V += np.linalg.det(X.T @ dN) * w[q]
    np.testing.assert_allclose(F.sum(), V, rtol=1e-13)


def test_t10_affine_vertex_and_midside_shares():
    # on a straight-sided quadratic tetrahedron int N_vertex = -V/20 and
    # int N_midside = V/5; TetRule(4) integrates quadratics exactly up to its
    # typed constants 0.13819660 / 0.58541020 (about 1e-8).
    conn, xyz = O.t10blockx_a(np.array([0.0, 1.0]), np.array([0.0, 1.0]), np.array([0.0, 1.0]))
    conn1 = conn[:1]
    used = np.unique(conn1)
    F = O.linform_dot(O.T10, conn1, xyz, 1, [1.0])
    X = xyz[conn1[0, :4] - 1]
    V = abs(np.linalg.det(np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]]))) / 6.0
    np.testing.assert_allclose(F[conn1[0, :4] - 1], -V / 20, rtol=1e-7)
    np.testing.assert_allclose(F[conn1[0, 4:] - 1], V / 5, rtol=1e-7)
    assert np.count_nonzero(F) == len(used)


def test_sysvec_dof_errors_match_the_reference_text():
    conn, xyz = O.block_mesh(O.H8, (1, 1, 1), L=1.0)
    dof = np.arange(1, 9, dtype=np.int64).reshape(8, 1)
    bad = dof.copy()
    bad[3, 0] = 0
    with pytest.raises(ValueError, match="Row degree of freedom < 1"):
        O.linform_dot(O.H8, conn, xyz, 1, [1.0], bad, 8)
    bad[3, 0] = 9
    with pytest.raises(ValueError, match="Row degree of freedom > size"):
        O.linform_dot(O.H8, conn, xyz, 1, [1.0], bad, 8)


def test_general_numbering_permutes_the_vector():
    conn, xyz = O.block_mesh(O.H8, (2, 3, 2), L=1.0, amp_frac=0.1, seed=3)
    nn = xyz.shape[0]
    F0 = O.linform_dot(O.H8, conn, xyz, 3, [1.0, 2.0, 3.0])
    perm = np.random.default_rng(1).permutation(nn * 3) + 1
    F1 = O.linform_dot(O.H8, conn, xyz, 3, [1.0, 2.0, 3.0], perm.reshape(nn, 3), nn * 3)
    np.testing.assert_array_equal(F1[perm - 1], F0)
