"""Pins for the oracle's free-then-fixed numbering and free-free block
(numberdofs!, src/FieldModule.jl:343-377; matrix_blocked_ff,
src/MatrixUtilityModule.jl:675-690): the numbering's structure, brute-force
agreement with the node-major matrix restricted to the free DOFs, and the
mechanics (a clamped face removes the six rigid-body modes)."""
import numpy as np
import scipy.sparse as sp

import oracle as O


def dense(cp, rv, nz, n):
    return sp.csc_matrix((nz, rv - 1, cp - 1), shape=(n, n)).toarray()


def clamped(xyz, ndn):
    fixed = np.zeros((xyz.shape[0], ndn), bool)
    fixed[np.abs(xyz[:, 0]) < 1e-12, :] = True  # the face x = 0
    return fixed


def test_numbering_free_first_in_entity_order():
    fixed = np.array([[False, True], [True, True], [False, False]])
    dof, nfree = O.numberdofs_ebc(fixed)
    assert nfree == 3
    assert dof.tolist() == [[1, 4], [5, 6], [2, 3]]
    dof2, _ = O.numberdofs_ebc(fixed, entperm=[3, 1, 2])
    # free pass over entities 3, 1, 2, then the fixed pass in the same order
    assert dof2.tolist() == [[3, 4], [5, 6], [1, 2]]


def test_free_block_equals_node_major_restriction_bitwise():
    conn, xyz = O.block_mesh(O.H8, (3, 2, 2), L=1.0, amp_frac=0.1, seed=5)
    D = O.iso_D(200e3, 0.3)
    fixed = clamped(xyz, 3)
    dof, nfree = O.numberdofs_ebc(fixed)
    n = xyz.shape[0] * 3
    Kfree_then_fixed = O.assemble_csc(O.H8, O.LIN_ELASTIC, conn, xyz, 3, D.ravel(), dof, n)
    Kff = dense(*O.matrix_blocked_ff(*Kfree_then_fixed, nfree), nfree)
    Knm = dense(*O.assemble_csc(O.H8, O.LIN_ELASTIC, conn, xyz, 3, D.ravel()), n)
    keep = ~fixed.ravel()
    assert np.array_equal(Kff, Knm[np.ix_(keep, keep)])  # same sums in the same element order


def test_clamping_removes_the_rigid_modes():
    conn, xyz = O.block_mesh(O.H8, (3, 2, 2), L=1.0, amp_frac=0.1, seed=5)
    D = O.iso_D(200e3, 0.3)
    n = xyz.shape[0] * 3
    K = dense(*O.assemble_csc(O.H8, O.LIN_ELASTIC, conn, xyz, 3, D.ravel()), n)
    w = np.linalg.eigvalsh(K)
    assert np.sum(np.abs(w) < 1e-8 * w.max()) == 6  # rigid translations and rotations
    fixed = clamped(xyz, 3)
    dof, nfree = O.numberdofs_ebc(fixed)
    Kff = dense(*O.matrix_blocked_ff(*O.assemble_csc(O.H8, O.LIN_ELASTIC, conn, xyz, 3, D.ravel(), dof, n),
                                     nfree), nfree)
    np.linalg.cholesky(Kff)  # positive definite
