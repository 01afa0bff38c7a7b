"""Elements that list a node twice (a collapsed edge).  The reference
assembles such an element like any other: every (local column, local row)
pair of its matrix lands in the global block of the two nodes, so the
block of a repeated node sums several local pairs of one element.  Checked
on the fused and the staged paths against the oracle (pattern bit-exact,
FP64 values within 1e-12 relative Frobenius), and the fused kernel's tiles
against the host builder byte for byte."""
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


def collapsed_mesh(n, which):
    conn, xyz = core.gen_h8block(1.0, 1.0, 1.0, *n)
    xyz = core.gen_perturb(xyz, [0, 0, 0], [1, 1, 1], 0.1 / max(n), 7)
    conn = conn.copy()
    for e in which:  # node 2 of the element takes node 1's place: edge 1-2 collapses
        conn[e, 1] = conn[e, 0]
    return conn, xyz


FORMS = [("elasticity", L.LIN_ELASTIC, 3, O.iso_D(200e3, 0.3)), ("diffusion", L.DIFFUSION_ISO, 1, np.array([1.7]))]


@pytest.mark.parametrize("which", [[0], [0, 5, 11, 17]])
@pytest.mark.parametrize("path", ["auto", "staged"])
def test_repeated_node_matches_oracle(ctx, which, path):
    conn, xyz = collapsed_mesh((4, 3, 3), which)
    for fname, form, ndn, coeff in FORMS:
        say("collapsed elements", which, "path", path, "form", fname)
        ocp, orv, onz = O.assemble_csc(O.H8, form, conn, xyz, ndn, np.ravel(coeff))
        mesh = core.DeviceMesh(ctx, L.H8, conn, xyz)
        mesh.set_dofs(ndn)
        mesh.set_numeric_path(path)
        M = mesh.bilform(form, L.RULE_GAUSS, 2, np.ravel(coeff))
        cp, rv, nz = M.csc_arrays(base=1)
        assert np.array_equal(cp, ocp) and np.array_equal(rv, orv)
        err = float(np.linalg.norm(nz - onz) / np.linalg.norm(onz))
        say("  rel Frobenius", err)
        assert err <= 1e-12
        M.close()
        if path == "auto":
            nt, diff, _ = mesh.tiles_compare(form, L.FP64)
            say("  tiles", nt, "first difference", diff)
            assert diff == -1
        mesh.close()


def fan_mesh(nfan):
    """H8 4^3 block whose first nfan elements all take node 1 as their
    first node: that node has 8 nfan > 256 (element, node) candidates, more
    than the adjacency pass sorts in registers."""
    conn, xyz = core.gen_h8block(1.0, 1.0, 1.0, 4, 4, 4)
    xyz = core.gen_perturb(xyz, [0, 0, 0], [1, 1, 1], 0.025, 9)
    conn = conn.copy()
    conn[:nfan, 0] = 1
    return conn, xyz


@pytest.mark.parametrize("case", ["collapsed", "fan"])
def test_repeated_node_served(ctx, case):
    """Through the assembly server: a new mesh's first assemblies take the
    staged path with contribution lists written by the adjacency pass (or,
    for a node with more than 256 candidates, merged per block), later
    re-assemblies the fused kernel.  Every result matches the oracle, and the
    staged ones equal the staged path of a plain mesh (lists merged per
    block) bit for bit."""
    conn, xyz = collapsed_mesh((4, 3, 3), [0, 5, 11, 17]) if case == "collapsed" else fan_mesh(34)
    for fname, form, ndn, coeff in FORMS:
        say("served", case, "form", fname)
        coeff = np.ravel(coeff)
        ocp, orv, onz = O.assemble_csc(O.H8, form, conn, xyz, ndn, coeff)
        mesh = core.DeviceMesh(ctx, L.H8, conn, xyz)
        mesh.set_dofs(ndn)
        mesh.set_numeric_path("staged")
        M = mesh.bilform(form, L.RULE_GAUSS, 2, coeff)
        plain = M.csc_arrays(base=1)
        M.close()
        mesh.close()
        srv = core.Server(0, workers=1, hbm_fraction=0.5)
        try:
            rid, mid = srv.submit_mesh(L.H8, conn, xyz, ndn, form, L.RULE_GAUSS, 2, coeff, L.FP64, keep=True)
            for k in range(6):  # assemblies 1-4 staged, 5 and 6 fused
                if k > 0:
                    rid = srv.submit_reassembly(mid, form, L.RULE_GAUSS, 2, coeff, L.FP64)
                st, _, _ = srv.wait(rid)
                assert st == 2, (k, st)
                cp, rv, nz = srv.result(rid).csc_arrays(1)
                srv.release(rid)
                assert np.array_equal(cp, ocp) and np.array_equal(rv, orv)
                err = float(np.linalg.norm(nz - onz) / np.linalg.norm(onz))
                say("  assembly", k + 1, "rel Frobenius", err)
                assert err <= 1e-12
                if k < 4:
                    assert np.array_equal(nz, plain[2]), k
        finally:
            srv.close()


@pytest.mark.parametrize("etype", ["h8", "t10"])
def test_served_staged_equals_plain_staged(ctx, etype):
    """Contribution lists from the adjacency sort (server) and from the
    per-block merge (plain mesh) give the same staged matrices, bit for bit,
    on regular H8 and T10 meshes, FP64 and mixed."""
    if etype == "h8":
        conn, xyz = core.gen_h8block(1.0, 1.0, 1.0, 6, 5, 4)
        xyz = core.gen_perturb(xyz, [0, 0, 0], [1, 1, 1], 0.1 / 6, 3)
        et, rule = L.H8, (L.RULE_GAUSS, 2)
    else:
        conn, xyz = core.gen_t10block(1.0, 1.0, 1.0, 3, 4, 3, 0.1 / 4, 0.03 / 4, 5)
        et, rule = L.T10, (L.RULE_TET, 4)
    for fname, form, ndn, coeff in FORMS:
        for prec in (L.FP64, L.MIXED):
            say("served vs plain staged", etype, fname, "precision", prec)
            coeff = np.ravel(coeff)
            mesh = core.DeviceMesh(ctx, et, conn, xyz)
            mesh.set_dofs(ndn)
            mesh.set_numeric_path("staged")
            M = mesh.bilform(form, rule[0], rule[1], coeff, precision=prec)
            plain = M.csc_arrays(base=1)
            M.close()
            mesh.close()
            srv = core.Server(0, workers=1, hbm_fraction=0.5)
            try:
                rid, _ = srv.submit_mesh(et, conn, xyz, ndn, form, rule[0], rule[1], coeff, prec)
                st, _, _ = srv.wait(rid)
                assert st == 2
                got = srv.result(rid).csc_arrays(1)
                srv.release(rid)
            finally:
                srv.close()
            for x, y in zip(got, plain):
                assert np.array_equal(x, y)
