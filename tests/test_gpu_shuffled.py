"""Node and element numberings with no geometric order: a block whose nodes
and elements are randomly renumbered.  The tiles (built in the Morton order
of the coordinates) then hold scattered node and element numbers, and every
block still sums its elements in ascending element number.  Fused, staged
and served paths against the oracle (pattern bit-exact, FP64 within 1e-12
relative Frobenius, mixed within 1e-6); device tiles byte-identical to the
host builder's."""
import sys

import numpy as np
import pytest

from feasm import _lib as L
from feasm import core
import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_fro(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def shuffled(etype, seed):
    rng = np.random.default_rng(seed)
    if etype == "h8":
        conn, xyz = core.gen_h8block(1.0, 1.0, 1.0, 8, 7, 6)
        xyz = core.gen_perturb(xyz, [0, 0, 0], [1, 1, 1], 0.1 / 8, seed)
        et, oet, rule = L.H8, O.H8, (L.RULE_GAUSS, 2)
    else:
        conn, xyz = core.gen_t10block(1.0, 1.0, 1.0, 4, 3, 4, 0.1 / 4, 0.03 / 4, seed)
        et, oet, rule = L.T10, O.T10, (L.RULE_TET, 4)
    nn = xyz.shape[0]
    newnum = rng.permutation(nn)  # old node k (0-based) -> new number newnum[k]
    xyz2 = np.empty_like(xyz)
    xyz2[newnum] = xyz
    conn2 = newnum[conn - 1] + 1
    conn2 = np.ascontiguousarray(conn2[rng.permutation(conn2.shape[0])])
    return conn2, xyz2, et, oet, rule


FORMS = [("elasticity", L.LIN_ELASTIC, 3, np.ravel(O.iso_D(200e3, 0.3))),
         ("diffusion", L.DIFFUSION_ISO, 1, np.array([0.8])),
         ("mass", L.DOT, 3, np.ravel(2.5 * np.eye(3)))]


@pytest.mark.parametrize("etype", ["h8", "t10"])
def test_shuffled_numbering(ctx, etype):
    conn, xyz, et, oet, rule = shuffled(etype, 11)
    for fname, form, ndn, coeff in FORMS:
        ocp, orv, onz = O.assemble_csc(oet, form, conn, xyz, ndn, coeff)
        for path in ("auto", "staged"):
            for prec, tol in ((L.FP64, 1e-12), (L.MIXED, 1e-6)):
                say("shuffled", etype, fname, path, "precision", prec)
                m = core.DeviceMesh(ctx, et, conn, xyz)
                m.set_dofs(ndn)
                m.set_numeric_path(path)
                M = m.bilform(form, rule[0], rule[1], coeff, precision=prec)
                cp, rv, nz = M.csc_arrays(1)
                assert np.array_equal(cp, ocp) and np.array_equal(rv, orv)
                assert rel_fro(nz, onz) <= tol
                M.close()
                if path == "auto":
                    nt, diff, _ = m.tiles_compare(form, prec)
                    assert diff == -1, (nt, diff)
                m.close()
    srv = core.Server(0, workers=2, hbm_fraction=0.5)
    try:
        fname, form, ndn, coeff = FORMS[0]
        ocp, orv, onz = O.assemble_csc(oet, form, conn, xyz, ndn, coeff)
        rid, _ = srv.submit_mesh(et, conn, xyz, ndn, form, rule[0], rule[1], coeff, L.MIXED)
        st, _, _ = srv.wait(rid)
        assert st == 2
        cp, rv, nz = srv.result(rid).csc_arrays(1)
        srv.release(rid)
        assert np.array_equal(cp, ocp) and np.array_equal(rv, orv)
        assert rel_fro(nz, onz) <= 1e-6
    finally:
        srv.close()
