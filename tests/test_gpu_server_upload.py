"""Assembly-server uploads of large meshes, whose connectivity and coordinates
are staged and copied in pieces shared out to helper threads: the result is
bitwise the same mesh's staged assembly outside the server, for borrowed and
for copied submissions, and a connectivity entry naming a missing node is
reported whichever thread stages its piece.  Every case runs through the C
ABI."""
import sys

import numpy as np
import pytest

import feasm
from feasm import _lib as L
from feasm import core

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def big_mesh():
    # 343,000 elements: six 2 MiB pieces of connectivity and five of
    # coordinates, enough for the helper threads
    n = 70
    conn, xyz = core.gen_h8block(1.0, 1.0, 1.0, n, n, n)
    xyz = core.gen_perturb(xyz, [0, 0, 0], [1, 1, 1], 0.1 / n, 3)
    return conn, xyz


def test_large_mesh_uploads_match_direct_assembly(ctx):
    conn, xyz = big_mesh()
    m = core.DeviceMesh(ctx, L.H8, conn, xyz)
    m.set_dofs(1)
    m.set_numeric_path("staged")
    M = m.bilform(L.DIFFUSION_ISO, L.RULE_GAUSS, 2, [1.7])
    ref = M.csc_arrays(1)
    M.close()
    m.close()
    srv = core.Server(0, workers=2, hbm_fraction=0.5)
    try:
        for borrow in (True, False):
            say("large mesh, borrowed" if borrow else "large mesh, copied")
            rid, _ = srv.submit_mesh(L.H8, conn, xyz, 1, L.DIFFUSION_ISO, L.RULE_GAUSS, 2, [1.7], L.FP64,
                                     borrow=borrow)
            st, _, _ = srv.wait(rid)
            assert st == 2
            got = srv.result(rid).csc_arrays(1)
            for x, y in zip(got, ref):
                assert np.array_equal(x, y)
            srv.release(rid)
    finally:
        srv.close()


@pytest.mark.parametrize("where", ["first", "last"])
def test_missing_node_in_any_piece_is_reported(where):
    conn, xyz = big_mesh()
    bad = conn.copy()
    bad.flat[0 if where == "first" else bad.size - 1] = xyz.shape[0] + 7
    say("bad connectivity entry in the", where, "piece")
    srv = core.Server(0, workers=1, hbm_fraction=0.5)
    try:
        rid, _ = srv.submit_mesh(L.H8, bad, xyz, 1, L.DIFFUSION_ISO, L.RULE_GAUSS, 2, [1.0], L.FP64, borrow=True)
        st, _, _ = srv.wait(rid)
        assert st != 2
        with pytest.raises(feasm.FeasmError, match="missing node"):
            srv.result(rid)
        srv.release(rid)
        # the server keeps working afterwards
        rid, _ = srv.submit_mesh(L.H8, conn, xyz, 1, L.DIFFUSION_ISO, L.RULE_GAUSS, 2, [1.0], L.FP64, borrow=True)
        st, _, _ = srv.wait(rid)
        assert st == 2
        srv.release(rid)
    finally:
        srv.close()
