"""Pins for the oracle's SysmatAssemblerSparseSymm restatement
(src/AssemblyModule.jl:342-583): the reference's 7x7 known answer, exact
agreement with the plain assembler's diagonal and rounding-level agreement
off the diagonal, and the zero-dropping of S + transpose(S)."""
import numpy as np
import scipy.sparse as sp

import oracle as O
from test_oracle_assembly import M1, M2


def lower_coo(m, dofs):
    """assemble!(::SysmatAssemblerSparseSymm): for j, for i = j:n (:494-523)."""
    I, J, V = [], [], []
    for j in range(len(dofs)):
        for i in range(j, len(dofs)):
            I.append(dofs[i])
            J.append(dofs[j])
            V.append(m[i, j])
    return I, J, V


def dense(cp, rv, nz, n):
    return sp.csc_matrix((nz, rv - 1, cp - 1), shape=(n, n)).toarray()


def test_kat_7x7_symm_assembler():
    # test/test_basics.jl:78-133: the same testA from SysmatAssemblerSparseSymm
    testA = np.array([
        [2.85928, 1.21875, 0.891063, 0.891614, 2.56958, 0.0, 0.0],
        [1.21875, 1.15515, 0.716396, 0.0714644, 1.56825, 0.0, 0.0],
        [0.891063, 0.716396, 0.936979, 0.0, 1.36026, 0.0, 0.0],
        [0.891614, 0.0714644, 0.0, 0.661253, 0.813892, 0.0, 0.0],
        [2.56958, 1.56825, 1.36026, 0.813892, 4.15934, 0.0, 0.0],
        [0.0] * 7, [0.0] * 7])
    a = lower_coo(M1.T @ M1, [5, 2, 1, 4])
    b = lower_coo(M2.T @ M2, [2, 3, 1, 5])
    A = dense(*O.symm_makematrix(a[0] + b[0], a[1] + b[1], a[2] + b[2], 7), 7)
    assert np.abs(testA - A).max() < 1e-5
    assert np.array_equal(A, A.T)


def test_symm_drops_zero_sums_and_halves_the_diagonal():
    # lower triples (2,1) = 1 and (1,2) = -1 (local order reversed in a second
    # element) cancel in S + S': no stored entry; the diagonal 3 stays 3
    cp, rv, nz = O.symm_makematrix([2, 2, 1, 2], [1, 2, 2, 2], [1.0, 3.0, -1.0, 0.0], 2)
    A = dense(cp, rv, nz, 2)
    assert np.array_equal(A, [[0.0, 0.0], [0.0, 3.0]])
    assert cp.tolist() == [1, 1, 2] and rv.tolist() == [2]


def test_symm_agrees_with_the_plain_assembler_on_meshes():
    for et, form, ndn, c, cells in ((O.H8, O.LIN_ELASTIC, 3, O.iso_D(200e3, 0.3), (3, 2, 2)),
                                    (O.T10, O.DIFFUSION_ISO, 1, [1.3], (2, 2, 1))):
        conn, xyz = O.block_mesh(et, cells, L=1.0, amp_frac=0.1, seed=6)
        n = xyz.shape[0] * ndn
        full = dense(*O.assemble_csc(et, form, conn, xyz, ndn, np.ravel(c)), n)
        sym = dense(*O.assemble_csc_symm(et, form, conn, xyz, ndn, np.ravel(c)), n)
        assert np.array_equal(np.diag(sym), np.diag(full))  # the same diagonal sums, halved exactly
        assert np.abs(sym - full).max() <= 1e-13 * np.abs(full).max()
        assert np.array_equal(sym, sym.T)
        assert np.array_equal(sym != 0, full != 0)  # no exact zeros on distorted meshes
