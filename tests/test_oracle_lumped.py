"""Pins for the oracle's HRZ lumping (bilform_dot into
SysmatAssemblerSparseHRZLumpingSymm, src/AssemblyModule.jl:943-1141): closed
forms on undistorted hexahedra, total-mass conservation on distorted H8 and
T10 meshes, and the reference's DOF errors."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O


def diag_of(cp, rv, nz, n):
    A = sp.csc_matrix((nz, rv - 1, cp - 1), shape=(n, n))
    assert (A - sp.diags(A.diagonal())).count_nonzero() == 0  # diagonal only
    return A.diagonal()


def test_h8_undistorted_lumped_mass_is_a_nodal_share():
    # on a cube element int N_a^2 = V/27 and the element mass is rho V, so
    # ffactor = 27/8 and every node gets rho V / 8 per incident element
    n, rho = 3, 7.8e-9
    conn, xyz = O.block_mesh(O.H8, (n, n, n), L=1.0)
    cp, rv, nz = O.hrz_lumped_csc(O.H8, conn, xyz, 3, (rho * np.eye(3)).ravel())
    d = diag_of(cp, rv, nz, xyz.shape[0] * 3).reshape(-1, 3)
    count = np.bincount(conn.ravel() - 1, minlength=xyz.shape[0])
    np.testing.assert_allclose(d, (count * rho / n ** 3 / 8)[:, None] * np.ones((1, 3)), rtol=1e-13)


@pytest.mark.parametrize("etype,cells", [(O.H8, (4, 3, 3)), (O.T10, (2, 3, 2))])
def test_lumping_conserves_the_total_mass(etype, cells):
    # em2 = sum of the element's entries = rho * ndn * (sum_q J w) with
    # sum_a N_a = 1; the lumped diagonal sums to it: rho * ndn * V, V = 1
    # (det J integrated exactly on trilinear hexahedra; for curved T10 the
    # rule's volume, checked through the consistent mass's total)
    rho = 2.7
    conn, xyz = O.block_mesh(etype, cells, L=1.0, amp_frac=0.1, seed=4)
    nd = xyz.shape[0] * 3
    cp, rv, nz = O.hrz_lumped_csc(etype, conn, xyz, 3, (rho * np.eye(3)).ravel())
    d = diag_of(cp, rv, nz, nd)
    ccp, crv, cnz = O.assemble_csc(etype, O.DOT, conn, xyz, 3, (rho * np.eye(3)).ravel())
    total = cnz.sum()
    np.testing.assert_allclose(d.sum(), total, rtol=1e-13)
    if etype == O.H8:
        np.testing.assert_allclose(total, 3 * rho, rtol=1e-12)
    assert np.all(d > 0)  # HRZ keeps every lumped mass positive, quadratic tets included


def test_lumped_dof_errors_match_the_reference_text():
    conn, xyz = O.block_mesh(O.H8, (1, 1, 1), L=1.0)
    dof = np.arange(1, 9, dtype=np.int64).reshape(8, 1)
    dof[2, 0] = 0
    with pytest.raises(ValueError, match="Column degree of freedom < 1"):
        O.hrz_lumped_csc(O.H8, conn, xyz, 1, [1.0], dof, 8)
    dof[2, 0] = 9
    with pytest.raises(ValueError, match="Column degree of freedom > size"):
        O.hrz_lumped_csc(O.H8, conn, xyz, 1, [1.0], dof, 8)
