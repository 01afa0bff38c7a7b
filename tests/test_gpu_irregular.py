"""The device path on irregular meshes read by the importers: the
reference's NASTRAN T10 cylinder (829 elements, test/cylinder.nas) and the
NAFEMS LE10 Abaqus T10 mesh (60 elements), against the oracle on the same
connectivity and coordinates.  Bar: pattern bit-exact, FP64 values within
1e-12 relative Frobenius, mixed precision within 1e-6."""
import os
import sys

import numpy as np
import pytest

from feasm import _lib as L
from feasm import core
from feasm import finetools as F
import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = {
    "cylinder": lambda: F.import_NASTRAN(os.path.join(HERE, "golden", "meshes", "cylinder.nas"),
                                         expectfixedformat=True),
    "le10": lambda: F.import_ABAQUS(os.path.join(HERE, "golden", "meshes", "problem_LE10NAFEMS_MST10.inp")),
}
FORMS = [("elasticity", L.LIN_ELASTIC, O.iso_D(70e3, 0.33)), ("mass", L.DOT, 2.7e-9 * np.eye(3)),
         ("general D", L.LIN_ELASTIC, O.iso_D(70e3, 0.33) + np.diag([0.0, 1e3, 0.0, 0.0, 0.0, 5e2]))]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_fro(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("name", list(MESHES))
@pytest.mark.parametrize("path", ["fused", "staged"])
def test_irregular_t10_meshes_match_oracle(ctx, name, path):
    out = MESHES[name]()
    conn, xyz = out["fesets"][0].conn, out["fens"].xyz
    for fname, form, coeff in FORMS:
        say("mesh", name, "path", path, "form", fname, "elements", conn.shape[0])
        ocp, orv, onz = O.assemble_csc(O.T10, form, conn, xyz, 3, np.ravel(coeff))
        mesh = core.DeviceMesh(ctx, L.T10, conn, xyz)
        mesh.set_dofs(3)
        mesh.set_numeric_path(path)
        for prec, tol in ((L.FP64, 1e-12), (L.MIXED, 1e-6)):
            M = mesh.bilform(form, L.RULE_TET, 4, np.ravel(coeff), precision=prec)
            cp, rv, nz = M.csc_arrays(base=1)
            assert np.array_equal(cp, ocp) and np.array_equal(rv, orv)
            err = rel_fro(nz, onz)
            say("  precision", prec, "rel Frobenius", err)
            assert err <= tol
            M.close()
        mesh.close()
