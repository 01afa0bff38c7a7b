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
