"""GPU bilform_* into a SysmatAssemblerSparseSymm (the full symmetric
assembly, then dropzeros on the device) against the oracle's restatement of
the symmetric assembler: pattern bit-exact (exact zeros dropped as
S + transpose(S) drops them), values within 1e-12 relative Frobenius."""
import sys

import numpy as np
import pytest

from feasm import _lib as L
from feasm import core
from feasm import finetools as F
import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name,form,ndn,coeff,amp", [
    ("H8 mass rho I3 (exact zeros off the component diagonal)", L.DOT, 3, 7.8e-9 * np.eye(3), 0.1),
    ("H8 elasticity", L.LIN_ELASTIC, 3, O.iso_D(200e3, 0.3), 0.15),
    ("H8 diffusion", L.DIFFUSION_ISO, 1, np.array([1.0]), 0.1),
])
def test_symm_assembler_matches_oracle(ctx, name, form, ndn, coeff, amp):
    say("case", name)
    n = (4, 3, 3)
    conn, xyz = core.gen_h8block(1.0, 1.0, 1.0, *n)
    xyz = core.gen_perturb(xyz, [0, 0, 0], [1, 1, 1], amp / max(n), 8)
    m = core.DeviceMesh(ctx, O.H8, conn, xyz)
    m.set_dofs(ndn)
    full = m.bilform(form, L.RULE_GAUSS, 2, coeff)
    sym = full.dropzeros()
    cp, rv, nz = sym.csc_arrays(1)
    ocp, orv, onz = O.assemble_csc_symm(O.H8, form, conn, xyz, ndn, np.ravel(coeff))
    assert np.array_equal(cp, ocp) and np.array_equal(rv, orv)
    err = np.linalg.norm(nz - onz) / np.linalg.norm(onz)
    say(name, "rel err", err, "nnz full", full.nnz, "nnz symm", sym.nnz)
    assert err <= 1e-12
    if form == L.DOT:
        assert sym.nnz * 3 == full.nnz  # only the component-diagonal entries survive
    full.close()
    sym.close()
    m.close()


def test_mirror_default_symmetric_assembler(ctx):
    fens, fes = F.H8block(1.0, 1.0, 2.0, 2, 2, 3)
    geom = F.NodalField(fens.xyz)
    u = F.NodalField(np.zeros((fens.count(), 3)))
    F.numberdofs(u)
    femm = F.FEMMBase(F.IntegDomain(fes, F.GaussRule(3, 2)), ctx=ctx)
    Ms = F.bilform_dot(femm, F.SysmatAssemblerSparseSymm(0.0, ctx=ctx), geom, u, F.DataCache(2.0 * np.eye(3)), 3)
    Mp = F.bilform_dot(femm, F.SysmatAssemblerSparse(0.0, ctx=ctx), geom, u, F.DataCache(2.0 * np.eye(3)), 3)
    assert Ms.nnz * 3 == Mp.nnz
    assert abs(Ms - Mp).max() == 0.0  # the same values, without the stored zeros
