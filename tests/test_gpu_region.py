"""Meshes with nodes of no element: a region of a block (only the elements
whose centroid lies in x < 0.45, over all the block's nodes, as when one
material region is assembled on the global node set) and a mesh with no
element at all.  The node-major CSC has empty columns for those nodes; the
fused, staged and served paths match the oracle (pattern bit-exact, FP64
values within 1e-12 relative Frobenius, mixed within 1e-6), and the device
tile builder matches the host builder byte for byte."""
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


def region(etype):
    if etype == "h8":
        conn, xyz = core.gen_h8block(1.0, 1.0, 1.0, 9, 7, 5)
        xyz = core.gen_perturb(xyz, [0, 0, 0], [1, 1, 1], 0.1 / 9, 6)
        et, oet, rule = L.H8, O.H8, (L.RULE_GAUSS, 2)
    else:
        conn, xyz = core.gen_t10block(1.0, 1.0, 1.0, 5, 3, 3, 0.1 / 5, 0.03 / 5, 6)
        et, oet, rule = L.T10, O.T10, (L.RULE_TET, 4)
    cx = xyz[conn - 1, 0].mean(axis=1)
    return np.ascontiguousarray(conn[cx < 0.45]), xyz, et, oet, rule


FORMS = [("elasticity", L.LIN_ELASTIC, 3, np.ravel(O.iso_D(200e3, 0.3))),
         ("diffusion", L.DIFFUSION_ISO, 1, np.array([1.3]))]


@pytest.mark.parametrize("etype", ["h8", "t10"])
def test_region_of_a_block(ctx, etype):
    conn, xyz, et, oet, rule = region(etype)
    unused = xyz.shape[0] - len(np.unique(conn))
    say("region", etype, "elements", conn.shape[0], "nodes", xyz.shape[0], "of no element", unused)
    assert unused > 0
    for fname, form, ndn, coeff in FORMS:
        ocp, orv, onz = O.assemble_csc(oet, form, conn, xyz, ndn, coeff)
        for path in ("auto", "staged"):
            for prec, tol in ((L.FP64, 1e-12), (L.MIXED, 1e-6)):
                say("  form", fname, "path", path, "precision", prec)
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
                    say("    tiles", nt, "first difference", diff)
                    assert diff == -1
                m.close()
        srv = core.Server(0, workers=1, hbm_fraction=0.5)
        try:
            rid, mid = srv.submit_mesh(et, conn, xyz, ndn, form, rule[0], rule[1], coeff, L.FP64, keep=True)
            for k in range(6):  # staged first, fused from the fifth assembly
                if k:
                    rid = srv.submit_reassembly(mid, form, rule[0], rule[1], coeff, L.FP64)
                st, _, _ = srv.wait(rid)
                assert st == 2, (k, st)
                cp, rv, nz = srv.result(rid).csc_arrays(1)
                srv.release(rid)
                assert np.array_equal(cp, ocp) and np.array_equal(rv, orv)
                assert rel_fro(nz, onz) <= 1e-12
        finally:
            srv.close()


def test_mesh_with_no_element(ctx):
    """Nodes only: every column is empty."""
    _, xyz = core.gen_h8block(1.0, 1.0, 1.0, 2, 2, 2)
    conn = np.zeros((0, 8), np.int64)
    m = core.DeviceMesh(ctx, L.H8, conn, xyz)
    m.set_dofs(3)
    M = m.bilform(L.LIN_ELASTIC, L.RULE_GAUSS, 2, np.ravel(O.iso_D(200e3, 0.3)))
    cp, rv, nz = M.csc_arrays(1)
    assert M.nnz == 0 and len(rv) == 0 and len(nz) == 0
    assert np.array_equal(cp, np.ones(3 * xyz.shape[0] + 1, np.int64))
    M.close()
    m.close()
