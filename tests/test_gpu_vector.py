"""GPU load vectors (linform_dot + SysvecAssembler) against the oracle.

Bar: within 1e-14 relative (the device performs the reference's additions in
the reference's order; only the Jacobian's association may differ), general
DOF numberings, the full-size C2 mesh through the total-load property, and
the error paths.  Every case runs through the C ABI."""
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


CASES = [
    ("H8 3 DOF", O.H8, (5, 4, 3), 0.15, 3, [1.0, -2.0, 0.5]),
    ("H8 1 DOF", O.H8, (6, 6, 6), 0.1, 1, [9.81]),
    ("T10 3 DOF", O.T10, (3, 3, 2), 0.1, 3, [0.0, 0.0, -2.7e-9]),
    ("T10 2 DOF", O.T10, (2, 3, 2), 0.1, 2, [1.5, -0.5]),
]


def mesh_for(etype, n, amp, seed):
    if etype == O.H8:
        conn, xyz = core.gen_h8block(1.0, 1.0, 1.0, *n)
        if amp:
            xyz = core.gen_perturb(xyz, [0, 0, 0], [1, 1, 1], amp / max(n), seed)
        return conn, xyz, (L.RULE_GAUSS, 2)
    h = 1.0 / max(n)
    conn, xyz = core.gen_t10block(1.0, 1.0, 1.0, *n, amp * h, 0.3 * amp * h, seed)
    return conn, xyz, (L.RULE_TET, 4)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_load_vector_matches_oracle(ctx, case):
    name, et, n, amp, ndn, f = case
    say("case", name)
    conn, xyz, rule = mesh_for(et, n, amp, 11)
    m = core.DeviceMesh(ctx, et, conn, xyz)
    m.set_dofs(ndn)
    got = m.linform_dot(rule[0], rule[1], f, xyz.shape[0] * ndn)
    want = O.linform_dot(et, conn, xyz, ndn, f)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    say(name, "rel err", err)
    assert err <= 1e-14
    again = m.linform_dot(rule[0], rule[1], f, xyz.shape[0] * ndn)
    assert np.array_equal(got, again)  # bitwise repeatable
    m.close()


RULES = [("H8 GaussRule(3,3)", O.H8, (4, 3, 3), 3, (L.RULE_GAUSS, 3)),
         ("H8 GaussRule(3,1)", O.H8, (4, 3, 3), 2, (L.RULE_GAUSS, 1)),
         ("T10 TetRule(5)", O.T10, (2, 2, 2), 3, (L.RULE_TET, 5)),
         ("T10 TetRule(1)", O.T10, (2, 2, 2), 1, (L.RULE_TET, 1))]


@pytest.mark.parametrize("case", RULES, ids=[c[0] for c in RULES])
def test_other_rules_take_the_element_vector_path(ctx, case):
    """Rules other than H8 2x2x2 and T10 4-point stage whole element vectors;
    same bar against the oracle run with the same rule."""
    name, et, n, ndn, rule = case
    say("case", name)
    conn, xyz, _ = mesh_for(et, n, 0.1, 13)
    m = core.DeviceMesh(ctx, et, conn, xyz)
    m.set_dofs(ndn)
    f = [1.0, -0.5, 2.0][:ndn]
    got = m.linform_dot(rule[0], rule[1], f, xyz.shape[0] * ndn)
    want = O.linform_dot(et, conn, xyz, ndn, f, rule=rule)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    say(name, "rel err", err)
    assert err <= 1e-14
    m.close()


def test_general_numbering_and_unused_dofs(ctx):
    conn, xyz, rule = mesh_for(O.H8, (3, 4, 2), 0.1, 5)
    nn = xyz.shape[0]
    rng = np.random.default_rng(2)
    nall = nn * 3 + 7  # seven DOFs no element touches stay zero
    perm = rng.permutation(nall)[: nn * 3] + 1
    dof = perm.reshape(nn, 3)
    m = core.DeviceMesh(ctx, O.H8, conn, xyz)
    m.set_dofs(3, dof, nall)
    got = m.linform_dot(rule[0], rule[1], [1.0, 2.0, 3.0], nall)
    want = O.linform_dot(O.H8, conn, xyz, 3, [1.0, 2.0, 3.0], dof, nall)
    assert np.linalg.norm(got - want) <= 1e-14 * np.linalg.norm(want)
    assert np.count_nonzero(got) == nn * 3
    m.close()


def test_full_size_c2_total_load(ctx):
    # H8 100^3 with the benchmark's perturbation: the total load is f * V = f
    n = 100
    conn, xyz = core.gen_h8block(1.0, 1.0, 1.0, n, n, n)
    xyz = core.gen_perturb(xyz, [0, 0, 0], [1, 1, 1], 0.15 / n, 2)
    m = core.DeviceMesh(ctx, O.H8, conn, xyz)
    m.set_dofs(3)
    f = np.array([0.0, -9.81, 2.0])
    got = m.linform_dot(L.RULE_GAUSS, 2, f, xyz.shape[0] * 3).reshape(-1, 3)
    np.testing.assert_allclose(got.sum(axis=0), f, rtol=1e-10, atol=1e-12)
    # interior nodes of the undistorted grid would carry 8 * V_e / 8 = h^3
    assert abs(np.median(got[:, 1]) / (-9.81 * 1e-6) - 1.0) < 0.2
    m.close()


def test_errors(ctx):
    conn, xyz, rule = mesh_for(O.H8, (2, 2, 2), 0.0, 0)
    m = core.DeviceMesh(ctx, O.H8, conn, xyz)
    m.set_dofs(3)
    with pytest.raises(feasm.FeasmError, match="one value per DOF"):
        m.linform_dot(rule[0], rule[1], [1.0, 2.0], xyz.shape[0] * 3)
    nn = xyz.shape[0]
    dup = np.arange(1, nn * 3 + 1, dtype=np.int64).reshape(nn, 3)
    dup[1, 0] = dup[0, 0]
    m.set_dofs(3, dup, nn * 3)
    with pytest.raises(feasm.FeasmError, match="distinct DOF numbers"):
        m.linform_dot(rule[0], rule[1], [1.0, 2.0, 3.0], nn * 3)
    m.close()


def test_mirror_linform_dot_and_distribloads(ctx):
    fens = F.H8block(2.0, 1.0, 1.0, 4, 2, 2)
    fes = fens[1]
    geom = F.NodalField(fens[0].xyz)
    u = F.NodalField(np.zeros((fens[0].count(), 3)))
    F.numberdofs(u)
    femm = F.FEMMBase(F.IntegDomain(fes, F.GaussRule(3, 2)), ctx=ctx)
    fi = F.ForceIntensity([0.0, 0.0, -1.0])
    vec = F.distribloads(femm, F.SysvecAssembler(0.0, ctx=ctx), geom, u, fi, 3)
    np.testing.assert_allclose(vec.reshape(-1, 3).sum(axis=0), [0.0, 0.0, -2.0], rtol=1e-13)
    a = F.SysvecAssembler(0.0, ctx=ctx)
    F.startassembly(a, 4)
    F.assemble(a, [1.0, 2.0], [2, 4])
    F.assemble(a, [0.5], [2])
    np.testing.assert_array_equal(F.makevector(a), [0.0, 1.5, 0.0, 2.0])
    with pytest.raises(ValueError, match="Row degree of freedom > size"):
        F.assemble(a, [1.0], [5])
