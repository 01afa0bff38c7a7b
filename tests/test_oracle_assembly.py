"""Pins for the oracle's assemble!/sparse() restatement and mesh generators,
from the reference's own known answers and from brute force."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O

M1 = np.array([[0.24406, 0.599773, 0.833404, 0.0420141],
               [0.786024, 0.00206713, 0.995379, 0.780298],
               [0.845816, 0.198459, 0.355149, 0.224996]])
M2 = np.array([[0.146618, 0.53471, 0.614342, 0.737833],
               [0.479719, 0.41354, 0.00760941, 0.836455],
               [0.254868, 0.476189, 0.460794, 0.00919633],
               [0.159064, 0.261821, 0.317078, 0.77646],
               [0.643538, 0.429817, 0.59788, 0.958909]])


def coo(m, rows, cols):
    """assemble!'s column-major serialization (src/AssemblyModule.jl:266-279)."""
    I, J, V = [], [], []
    for j in range(len(cols)):
        for i in range(len(rows)):
            I.append(rows[i])
            J.append(cols[j])
            V.append(m[i, j])
    return I, J, V


def dense(cp, rv, nz, m, n):
    return sp.csc_matrix((nz, rv - 1, cp - 1), shape=(m, n)).toarray()


def test_kat_7x7_symmetric():
    # test/test_basics.jl:78-133 (mmassembly2ya1)
    testA = np.array([
        [2.85928, 1.21875, 0.891063, 0.891614, 2.56958, 0.0, 0.0],
        [1.21875, 1.15515, 0.716396, 0.0714644, 1.56825, 0.0, 0.0],
        [0.891063, 0.716396, 0.936979, 0.0, 1.36026, 0.0, 0.0],
        [0.891614, 0.0714644, 0.0, 0.661253, 0.813892, 0.0, 0.0],
        [2.56958, 1.56825, 1.36026, 0.813892, 4.15934, 0.0, 0.0],
        [0.0] * 7, [0.0] * 7])
    a = coo(M1.T @ M1, [5, 2, 1, 4], [5, 2, 1, 4])
    b = coo(M2.T @ M2, [2, 3, 1, 5], [2, 3, 1, 5])
    A = dense(*O.sparse(a[0] + b[0], a[1] + b[1], a[2] + b[2], 7, 7), 7, 7)
    assert np.abs(testA - A).max() < 1e-5


def test_kat_nomatrixresult_entries_rectangular():
    # test/test_basics.jl:1577-1613 (massnomat1): rectangular element matrices
    a = coo(M1, [1, 7, 5], [5, 2, 1, 4])
    b = coo(M2, [2, 3, 1, 7, 5], [6, 7, 3, 4])
    A = dense(*O.sparse(a[0] + b[0], a[1] + b[1], a[2] + b[2], 7, 7), 7, 7)
    assert abs(A[0, 0] - 0.833404) < 1e-12
    assert abs(A[4, 0] - 0.355149) < 1e-12
    assert abs(A[6, 5] - 0.159064) < 1e-12
    assert abs(A[2, 6] - 0.41354) < 1e-12
    assert abs(A[6, 6] - 0.261821) < 1e-12
    # test/test_basics.jl:1464-1499 (mmassembl2): against a dense brute force
    ref = np.zeros((7, 7))
    for m, r, c in ((M1, [1, 7, 5], [5, 2, 1, 4]), (M2, [2, 3, 1, 4, 5], [6, 7, 3, 4])):
        for j, cc in enumerate(c):
            for i, rr in enumerate(r):
                ref[rr - 1, cc - 1] += m[i, j]
    a = coo(M1, [1, 7, 5], [5, 2, 1, 4])
    b = coo(M2, [2, 3, 1, 4, 5], [6, 7, 3, 4])
    A = dense(*O.sparse(a[0] + b[0], a[1] + b[1], a[2] + b[2], 7, 7), 7, 7)
    assert np.abs(ref - A).max() < 1e-15


def test_sparse_left_fold_in_input_order_and_explicit_zeros():
    # duplicates fold left to right in input order: (1e16 + 1) - 1e16 == 0
    cp, rv, nz = O.sparse([2, 2, 2, 1], [1, 1, 1, 3], [1e16, 1.0, -1e16, 0.0], 3, 3)
    assert cp.tolist() == [1, 2, 2, 3]
    assert rv.tolist() == [2, 1]
    assert nz[0] == 0.0            # not 1.0 (another order) -- the fold order is pinned
    assert nz[1] == 0.0            # explicit zero kept as a stored entry
    cp, rv, nz = O.sparse([2, 2, 2], [1, 1, 1], [1.0, 1e16, -1e16], 3, 3)
    assert nz[0] == 0.0           # (1 + 1e16) - 1e16 = 0 as well
    cp, rv, nz = O.sparse([2, 2, 2], [1, 1, 1], [1e16, -1e16, 1.0], 3, 3)
    assert nz[0] == 1.0


def test_sparse_random_vs_dense_sorted_rows():
    rng = np.random.default_rng(0)
    m, n, t = 37, 29, 2000
    I = rng.integers(1, m + 1, t)
    J = rng.integers(1, n + 1, t)
    V = rng.normal(size=t)
    cp, rv, nz = O.sparse(I, J, V, m, n)
    ref = np.zeros((m, n))
    np.add.at(ref, (I - 1, J - 1), V)
    assert np.allclose(dense(cp, rv, nz, m, n), ref, atol=1e-12)
    for j in range(n):
        col = rv[cp[j] - 1:cp[j + 1] - 1]
        assert np.all(np.diff(col) > 0)
    assert len(nz) == len(set(zip(I.tolist(), J.tolist())))
    with pytest.raises(ValueError):
        O.sparse([0], [1], [1.0], 3, 3)


def test_h8block_order_closed_form():
    # src/MeshHexahedronModule.jl:76-105
    nL, nW, nH = 3, 4, 2
    conn, xyz = O.h8block(3.0, 4.0, 2.0, nL, nW, nH)
    assert conn.shape == (24, 8) and xyz.shape == (60, 3)
    assert xyz[1].tolist() == [1.0, 0.0, 0.0] and xyz[nL + 1].tolist() == [0.0, 1.0, 0.0]
    assert conn[0].tolist() == [1, 2, 6, 5, 21, 22, 26, 25]
    assert conn[1].tolist() == [21, 22, 26, 25, 41, 42, 46, 45]  # z fastest
    assert conn[nH].tolist()[0] == 5                                  # then y


def test_h8_degree_and_nnz_closed_form():
    # test/test_miscellaneous.jl:931-952: node degree max 27 / min 8;
    # nnz of the node graph of an n^3 block is (3n+1)^3
    conn, xyz = O.block_mesh(O.H8, (3, 4, 5))
    cp, rv, nz = O.assemble_csc(O.H8, O.DIFFUSION_ISO, conn, xyz, 1, [1.0])
    deg = np.diff(cp)
    assert deg.max() == 27 and deg.min() == 8
    assert len(rv) == (3 * 3 + 1) * (3 * 4 + 1) * (3 * 5 + 1)
    cp3, rv3, _ = O.assemble_csc(O.H8, O.LIN_ELASTIC, conn, xyz, 3, O.iso_D(1.0, 0.3).ravel())
    assert len(rv3) == 9 * len(rv)


def test_t10blockx_counts():
    # test/test_meshing.jl:1458-1475: 5x10x5 -> 2541 nodes, 1500 elements
    xs = O.linearspace(0.0, 1.0, 6)
    ys = O.linearspace(0.0, 1.0, 11)
    conn, xyz = O.t10blockx_a(xs, ys, xs)
    assert conn.shape == (1500, 10) and xyz.shape == (2541, 3)
    # mid-edge node of the first element's first edge is numbered first
    assert conn[0, 4] == 6 * 11 * 6 + 1
    e = conn[0]
    assert np.allclose(xyz[e[4] - 1], (xyz[e[0] - 1] + xyz[e[1] - 1]) / 2)
    # every node used, positive volumes
    assert len(np.unique(conn)) == xyz.shape[0]
    assert O.min_detj(O.T10, conn, xyz) > 0


def test_determinism_repeat_assembly():
    # test/test_basics.jl:3016-3051: successive assemblies match exactly
    conn, xyz = O.block_mesh(O.H8, (3, 3, 3), amp_frac=0.15, seed=1)
    D = O.iso_D(200e3, 0.3).ravel()
    r1 = O.assemble_csc(O.H8, O.LIN_ELASTIC, conn, xyz, 3, D)
    r2 = O.assemble_csc(O.H8, O.LIN_ELASTIC, conn, xyz, 3, D)
    for x, y in zip(r1, r2):
        assert np.array_equal(x, y)


def test_dof_bounds_errors_like_assemble():
    conn, xyz = O.block_mesh(O.H8, (1, 1, 1))
    d = np.arange(1, 9, dtype=np.int64).reshape(8, 1)
    d[3, 0] = 0
    with pytest.raises(ValueError, match="< 1"):
        O.bilform_coo(O.H8, O.DIFFUSION_ISO, conn, xyz, 1, d, 8, [1.0])
    d[3, 0] = 9
    with pytest.raises(ValueError, match="> size"):
        O.bilform_coo(O.H8, O.DIFFUSION_ISO, conn, xyz, 1, d, 8, [1.0])


def test_perturbation_keeps_boundary_and_is_seeded():
    conn, xyz0 = O.block_mesh(O.H8, (4, 4, 4))
    x1 = O.perturb(xyz0, [0, 0, 0], [1, 1, 1], 0.025, 42)
    x2 = O.perturb(xyz0, [0, 0, 0], [1, 1, 1], 0.025, 42)
    assert np.array_equal(x1, x2)
    onb = np.any((xyz0 == 0) | (xyz0 == 1), axis=1)
    assert np.array_equal(x1[onb], xyz0[onb])
    assert np.all(np.abs(x1 - xyz0) <= 0.025) and np.any(x1[~onb] != xyz0[~onb])
