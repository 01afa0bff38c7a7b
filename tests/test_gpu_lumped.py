"""GPU HRZ-lumped mass (bilform_dot into SysmatAssemblerSparseHRZLumpingSymm)
against the oracle: the diagonal pattern bit-exact, values within 1e-12
relative Frobenius, general DOF numberings, the total-mass property at the
C2 size, and the error paths.  Every case runs through the C ABI."""
import sys

import numpy as np
import pytest

import feasm
from feasm import _lib as L
from feasm import core
from feasm import finetools as F
import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mesh_for(etype, n, amp, seed):
    if etype == O.H8:
        conn, xyz = core.gen_h8block(1.0, 1.0, 1.0, *n)
        if amp:
            xyz = core.gen_perturb(xyz, [0, 0, 0], [1, 1, 1], amp / max(n), seed)
        return conn, xyz, (L.RULE_GAUSS, 2)
    h = 1.0 / max(n)
    conn, xyz = core.gen_t10block(1.0, 1.0, 1.0, *n, amp * h, 0.3 * amp * h, seed)
    return conn, xyz, (L.RULE_TET, 4)


CASES = [
    ("H8 rho I3", O.H8, (5, 4, 3), 0.15, 3, 7.8e-9 * np.eye(3)),
    ("H8 scalar", O.H8, (4, 4, 4), 0.1, 1, np.array([[2.5]])),
    ("T10 rho I3", O.T10, (3, 2, 3), 0.1, 3, 2.7e-9 * np.eye(3)),
    ("T10 anisotropic 2", O.T10, (2, 2, 2), 0.1, 2, np.array([[2.0, 0.5], [0.5, 1.0]])),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_lumped_mass_matches_oracle(ctx, case):
    name, et, n, amp, ndn, c = case
    say("case", name)
    conn, xyz, rule = mesh_for(et, n, amp, 9)
    m = core.DeviceMesh(ctx, et, conn, xyz)
    m.set_dofs(ndn)
    M = m.bilform_dot_lumped(rule[0], rule[1], c)
    cp, rv, nz = M.csc_arrays(1)
    ocp, orv, onz = O.hrz_lumped_csc(et, conn, xyz, ndn, c.ravel())
    assert np.array_equal(cp, ocp) and np.array_equal(rv, orv)
    err = np.linalg.norm(nz - onz) / np.linalg.norm(onz)
    say(name, "rel err", err)
    assert err <= 1e-12
    M.close()
    m.close()


RULE_CASES = [
    ("H8 Gauss 1", O.H8, (3, 3, 2), 3, (L.RULE_GAUSS, 1)),
    ("H8 Gauss 3", O.H8, (4, 3, 3), 3, (L.RULE_GAUSS, 3)),   # 27 points: 64-thread blocks
    ("H8 Gauss 4", O.H8, (3, 3, 3), 2, (L.RULE_GAUSS, 4)),   # 64 points, the largest rule
    ("T10 Tet 1", O.T10, (2, 2, 2), 3, (L.RULE_TET, 1)),
    ("T10 Tet 5", O.T10, (2, 3, 2), 1, (L.RULE_TET, 5)),     # a negative weight
]


@pytest.mark.parametrize("case", RULE_CASES, ids=[c[0] for c in RULE_CASES])
def test_lumped_mass_other_rules(ctx, case):
    name, et, n, ndn, rule = case
    say("case", name)
    conn, xyz, _ = mesh_for(et, n, 0.1, 5)
    c = np.diag(np.arange(1.0, ndn + 1.0)) + 0.25 * (np.ones((ndn, ndn)) - np.eye(ndn))
    m = core.DeviceMesh(ctx, et, conn, xyz)
    m.set_dofs(ndn)
    M = m.bilform_dot_lumped(rule[0], rule[1], c)
    cp, rv, nz = M.csc_arrays(1)
    ocp, orv, onz = O.hrz_lumped_csc(et, conn, xyz, ndn, c.ravel(), rule=rule)
    assert np.array_equal(cp, ocp) and np.array_equal(rv, orv)
    err = np.linalg.norm(nz - onz) / np.linalg.norm(onz)
    say(name, "rel err", err)
    assert err <= 1e-12
    M.close()
    m.close()


def test_general_numbering_and_untouched_dofs(ctx):
    conn, xyz, rule = mesh_for(O.H8, (3, 2, 2), 0.1, 3)
    nn = xyz.shape[0]
    nall = nn * 3 + 5
    dof = (np.random.default_rng(4).permutation(nall)[: nn * 3] + 1).reshape(nn, 3)
    m = core.DeviceMesh(ctx, O.H8, conn, xyz)
    m.set_dofs(3, dof, nall)
    M = m.bilform_dot_lumped(rule[0], rule[1], 1.5 * np.eye(3))
    cp, rv, nz = M.csc_arrays(1)
    ocp, orv, onz = O.hrz_lumped_csc(O.H8, conn, xyz, 3, (1.5 * np.eye(3)).ravel(), dof, nall)
    assert np.array_equal(cp, ocp) and np.array_equal(rv, orv)
    assert len(nz) == nn * 3
    assert np.linalg.norm(nz - onz) <= 1e-12 * np.linalg.norm(onz)
    M.close()
    m.close()


def test_c2_size_total_mass(ctx):
    n = 100
    conn, xyz = core.gen_h8block(1.0, 1.0, 1.0, n, n, n)
    xyz = core.gen_perturb(xyz, [0, 0, 0], [1, 1, 1], 0.15 / n, 2)
    m = core.DeviceMesh(ctx, O.H8, conn, xyz)
    m.set_dofs(3)
    M = m.bilform_dot_lumped(L.RULE_GAUSS, 2, 7.8e-9 * np.eye(3))
    cp, rv, nz = M.csc_arrays(1)
    assert len(nz) == xyz.shape[0] * 3 and np.all(nz > 0)
    np.testing.assert_allclose(nz.sum(), 3 * 7.8e-9, rtol=1e-10)
    M.close()
    m.close()


def test_t10_full_size_total_mass(ctx):
    """T10 60^3 x 6 (1.3 M elements): every DOF gets a positive mass and the
    total is ndn * rho * volume (HRZ keeps each element's total, and the
    4-point rule integrates det J of these straight-sided tetrahedra exactly)."""
    n = 60
    conn, xyz = core.gen_t10block(1.0, 1.0, 1.0, n, n, n, 0.0, 0.0, 1)
    m = core.DeviceMesh(ctx, O.T10, conn, xyz)
    m.set_dofs(3)
    rho = 2.7e-9
    M = m.bilform_dot_lumped(L.RULE_TET, 4, rho * np.eye(3))
    cp, rv, nz = M.csc_arrays(1)
    assert len(nz) == xyz.shape[0] * 3 and np.all(nz > 0)
    assert np.array_equal(rv, np.arange(1, len(nz) + 1))
    np.testing.assert_allclose(nz.sum(), 3 * rho, rtol=1e-10)
    M.close()
    m.close()


def test_errors_and_mirror(ctx):
    conn, xyz, rule = mesh_for(O.H8, (2, 2, 2), 0.0, 0)
    m = core.DeviceMesh(ctx, O.H8, conn, xyz)
    m.set_dofs(3)
    with pytest.raises(feasm.FeasmError, match="ndn x ndn"):
        m.bilform_dot_lumped(rule[0], rule[1], [1.0, 2.0])
    m.close()
    fens, fes = F.H8block(1.0, 2.0, 1.0, 2, 3, 2)
    geom = F.NodalField(fens.xyz)
    u = F.NodalField(np.zeros((fens.count(), 3)))
    F.numberdofs(u)
    femm = F.FEMMBase(F.IntegDomain(fes, F.GaussRule(3, 2)), ctx=ctx)
    A = F.bilform_dot(femm, F.SysmatAssemblerSparseHRZLumpingSymm(0.0, ctx=ctx), geom, u,
                      F.DataCache(2.0 * np.eye(3)), 3)
    d = A.diagonal()
    assert A.nnz == len(d) and np.all(d > 0)
    np.testing.assert_allclose(d.sum(), 3 * 2.0 * 2.0, rtol=1e-13)  # ndn * rho * volume
