"""GPU free-free blocks (SysmatAssemblerFFBlock / matrix_blocked_ff) against
the oracle: a clamped face on H8 elasticity, the resident node-major
assembly restricted on the device to the free DOFs, bit-exact against the
oracle's assembly in the free-then-fixed numbering; the general (permuted)
numbering through the COO path; arbitrary masks against host slicing."""
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


def test_clamped_elasticity_free_block_matches_oracle(ctx):
    fens, fes = F.H8block(2.0, 1.0, 1.0, 6, 3, 3)
    xyz = core.gen_perturb(fens.xyz, [0, 0, 0], [2, 1, 1], 0.1 / 6, 3)
    geom = F.NodalField(xyz)
    u = F.NodalField(np.zeros((fens.count(), 3)))
    x0 = np.nonzero(np.abs(xyz[:, 0]) < 1e-12)[0] + 1
    for comp in (1, 2, 3):
        F.setebc(u, x0, True, comp, 0.0)
    F.numberdofs(u)
    nf = F.nfreedofs(u)
    femm = F.FEMMBase(F.IntegDomain(fes, F.GaussRule(3, 2)), ctx=ctx)
    D = O.iso_D(200e3, 0.3)
    Kff = F.bilform_lin_elastic(femm, F.SysmatAssemblerFFBlock(nf, ctx=ctx), geom, u, F.DeforModelRed3D,
                                F.DataCache(D))
    ocp, orv, onz = O.matrix_blocked_ff(*O.assemble_csc(O.H8, O.LIN_ELASTIC, fes.conn, xyz, 3, D.ravel(), u.dofnums,
                                                        F.nalldofs(u)), nf)
    say("nfree", nf, "nnz", Kff.nnz)
    assert Kff.shape == (nf, nf)
    assert np.array_equal(Kff.indptr, ocp - 1) and np.array_equal(Kff.indices, orv - 1)
    assert np.linalg.norm(Kff.data - onz) <= 1e-12 * np.linalg.norm(onz)
    # and bitwise the device's own node-major matrix restricted to the free DOFs
    full = F.bilform_lin_elastic(femm, F.SysmatAssemblerSparse(0.0, ctx=ctx), geom, u, F.DeforModelRed3D,
                                 F.DataCache(D))  # in the field's free-then-fixed numbering (COO path)
    assert np.linalg.norm((full[:nf, :nf] - Kff).toarray()) <= 1e-12 * np.linalg.norm(onz)
    # permuted entities: the field's own numbering through the COO path
    perm = np.random.default_rng(2).permutation(fens.count()) + 1
    F.numberdofs(u, perm)
    Kp = F.bilform_lin_elastic(femm, F.SysmatAssemblerFFBlock(nf, ctx=ctx), geom, u, F.DeforModelRed3D,
                               F.DataCache(D))
    pcp, prv, pnz = O.matrix_blocked_ff(*O.assemble_csc(O.H8, O.LIN_ELASTIC, fes.conn, xyz, 3, D.ravel(), u.dofnums,
                                                        F.nalldofs(u)), nf)
    assert np.array_equal(Kp.indptr, pcp - 1) and np.array_equal(Kp.indices, prv - 1)
    assert np.linalg.norm(Kp.data - pnz) <= 1e-12 * np.linalg.norm(pnz)


def test_free_block_of_arbitrary_masks(ctx):
    conn, xyz = core.gen_h8block(1.0, 1.0, 1.0, 4, 3, 5)
    m = core.DeviceMesh(ctx, O.H8, conn, xyz)
    m.set_dofs(3)
    K = m.bilform(L.LIN_ELASTIC, L.RULE_GAUSS, 2, O.iso_D(70e3, 0.33))
    cp, rv, nz = K.csc_arrays(0)
    n = len(cp) - 1
    rng = np.random.default_rng(7)
    for frac in (0.0, 0.3, 0.9, 1.0):
        keep = rng.random(n) < frac
        newidx = np.cumsum(keep) - 1
        rcp, rrv, rnz = [0], [], []
        for c in np.nonzero(keep)[0]:  # host restriction, stored zeros kept
            for k in range(cp[c], cp[c + 1]):
                if keep[rv[k]]:
                    rrv.append(newidx[rv[k]])
                    rnz.append(nz[k])
            rcp.append(len(rrv))
        B = K.free_block(keep)
        gcp, grv, gnz = B.csc_arrays(0)
        assert np.array_equal(gcp, rcp) and np.array_equal(grv, rrv)
        assert np.array_equal(gnz, rnz)
        B.close()
    K.close()
    m.close()
