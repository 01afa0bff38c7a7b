"""C2 at full size (H8 100^3, 3 DOF/node, FP64): the device matrix through
size-independent properties -- the pattern size ndn^2 (3n+1)^3, bitwise
symmetry (K x == K^T x), rigid-body modes in the null space, and sampled
columns against the oracle assembled on just the elements around them (a
column only sees the elements that hold its node, in ascending order) -- and
the mixed-precision error against FP64 at this size."""
import sys

import numpy as np
import pytest

from feasm import _lib as L
from feasm import core
import oracle as O

pytestmark = pytest.mark.gpu

N = 100


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c2(ctx):
    conn, xyz = core.gen_h8block(1.0, 1.0, 1.0, N, N, N)
    xyz = core.gen_perturb(xyz, [0, 0, 0], [1, 1, 1], 0.15 / N, 2)
    D = O.iso_D(200e3, 0.3)
    say("C2 full size: assembling", conn.shape[0], "elements")
    mesh = core.DeviceMesh(ctx, L.H8, conn, xyz)
    mesh.set_dofs(3)
    K = mesh.bilform(L.LIN_ELASTIC, L.RULE_GAUSS, 2, D.ravel())
    cp, rv, nz = K.csc_arrays(base=0)
    yield {"conn": conn, "xyz": xyz, "D": D, "mesh": mesh, "K": K, "cp": cp, "rv": rv, "nz": nz}
    K.close()
    mesh.close()


def test_pattern_size_and_symmetry(c2):
    import scipy.sparse as sp

    cp, rv, nz = c2["cp"], c2["rv"], c2["nz"]
    n = 3 * (N + 1) ** 3
    assert len(cp) == n + 1 and cp[-1] == 9 * (3 * N + 1) ** 3 == len(nz)
    assert np.all(np.diff(cp) > 0)
    # bitwise symmetry on 20000 sampled stored entries: K[j, i] is stored and equal
    rng = np.random.default_rng(3)
    idx = rng.integers(0, len(nz), 20000)
    cols = np.searchsorted(cp, idx, side="right") - 1
    rows = rv[idx]
    for p_, i, j in zip(idx, rows, cols):
        lo, hi = cp[i], cp[i + 1]
        q = lo + np.searchsorted(rv[lo:hi], j)
        assert q < hi and rv[q] == j and nz[q] == nz[p_]
    # and as an operator
    K = sp.csc_matrix((nz, rv, cp), shape=(n, n))
    x = rng.standard_normal(n)
    say("symmetry check: K x vs K^T x")
    y1, y2 = K @ x, K.T @ x
    assert np.linalg.norm(y1 - y2) <= 1e-13 * np.linalg.norm(y1)
    assert np.all(K.diagonal() > 0)


def test_rigid_body_modes(c2):
    import scipy.sparse as sp

    cp, rv, nz, xyz = c2["cp"], c2["rv"], c2["nz"], c2["xyz"]
    n = 3 * xyz.shape[0]
    K = sp.csc_matrix((nz, rv, cp), shape=(n, n))
    kmax = np.abs(nz).max()
    c = xyz - xyz.mean(axis=0)
    modes = []
    for k in range(3):  # translations
        u = np.zeros((xyz.shape[0], 3))
        u[:, k] = 1.0
        modes.append(u)
    for k in range(3):  # infinitesimal rotations about the axes
        w = np.zeros(3)
        w[k] = 1.0
        modes.append(np.cross(w, c))
    for u in modes:
        r = K @ u.ravel()
        rel = np.abs(r).max() / (kmax * np.abs(u).max())
        say("rigid-body residual", rel)
        assert rel <= 1e-10


def test_sampled_columns_match_oracle(c2):
    conn, xyz, D, cp, rv, nz = c2["conn"], c2["xyz"], c2["D"], c2["cp"], c2["rv"], c2["nz"]
    rng = np.random.default_rng(11)
    nodes = np.unique(np.concatenate([rng.choice(xyz.shape[0], 36, replace=False), [0, xyz.shape[0] - 1,
                                                                                  (N + 1) ** 2 // 2]]))
    near = np.nonzero(np.isin(conn, nodes + 1).any(axis=1))[0]  # elements holding a sampled node, in order
    say("sampled columns of", len(nodes), "nodes from", len(near), "elements")
    ocp, orv, onz = O.assemble_csc(O.H8, L.LIN_ELASTIC, conn[near], xyz, 3, D.ravel())
    for node in nodes:
        for k in range(3):
            j = 3 * node + k
            got_r, got_v = rv[cp[j]:cp[j + 1]], nz[cp[j]:cp[j + 1]]
            want_r, want_v = orv[ocp[j] - 1:ocp[j + 1] - 1] - 1, onz[ocp[j] - 1:ocp[j + 1] - 1]
            assert np.array_equal(got_r, want_r)
            err = np.linalg.norm(got_v - want_v) / np.linalg.norm(want_v)
            assert err <= 1e-12, (node, k, err)


def test_mixed_precision_error_at_full_size(c2):
    mesh, D, nz = c2["mesh"], c2["D"], c2["nz"]
    say("mixed precision at full size")
    M = mesh.bilform(L.LIN_ELASTIC, L.RULE_GAUSS, 2, D.ravel(), precision=L.MIXED)
    _, _, mz = M.csc_arrays(base=0)
    M.close()
    err = np.linalg.norm(mz - nz) / np.linalg.norm(nz)
    say("mixed rel Frobenius", err)
    assert err <= 1e-6


# ---------------------------------------------------------------- C3: T10 60^3 x 6
M3 = 60


@pytest.fixture(scope="module")
def c3(ctx):
    h = 1.0 / M3
    conn, xyz = core.gen_t10block(1.0, 1.0, 1.0, M3, M3, M3, 0.1 * h, 0.03 * h, 3)
    D = O.iso_D(70e3, 0.33)
    say("C3 full size: assembling", conn.shape[0], "T10 elements")
    mesh = core.DeviceMesh(ctx, L.T10, conn, xyz)
    mesh.set_dofs(3)
    yield {"conn": conn, "xyz": xyz, "D": D, "mesh": mesh}
    mesh.close()


def test_c3_stiffness_rigid_modes_and_sampled_columns(c3):
    import scipy.sparse as sp

    conn, xyz, D, mesh = c3["conn"], c3["xyz"], c3["D"], c3["mesh"]
    K = mesh.bilform(L.LIN_ELASTIC, L.RULE_TET, 4, D.ravel())
    cp, rv, nz = K.csc_arrays(base=0)
    K.close()
    n = 3 * xyz.shape[0]
    assert len(cp) == n + 1 and cp[-1] == len(nz) == 451604169
    A = sp.csc_matrix((nz, rv, cp), shape=(n, n))
    kmax = np.abs(nz).max()
    c = xyz - xyz.mean(axis=0)
    for k in range(3):
        u = np.zeros((xyz.shape[0], 3))
        u[:, k] = 1.0
        w = np.zeros(3)
        w[k] = 1.0
        for mode in (u, np.cross(w, c)):
            rel = np.abs(A @ mode.ravel()).max() / (kmax * np.abs(mode).max())
            say("C3 rigid-body residual", rel)
            assert rel <= 1e-10
    del A
    rng = np.random.default_rng(5)
    nodes = np.unique(rng.choice(xyz.shape[0], 30, replace=False))
    near = np.nonzero(np.isin(conn, nodes + 1).any(axis=1))[0]
    say("C3 sampled columns of", len(nodes), "nodes from", len(near), "elements")
    ocp, orv, onz = O.assemble_csc(O.T10, L.LIN_ELASTIC, conn[near], xyz, 3, D.ravel())
    for node in nodes:
        for k in range(3):
            j = 3 * node + k
            assert np.array_equal(rv[cp[j]:cp[j + 1]], orv[ocp[j] - 1:ocp[j + 1] - 1] - 1)
            want = onz[ocp[j] - 1:ocp[j + 1] - 1]
            assert np.linalg.norm(nz[cp[j]:cp[j + 1]] - want) <= 1e-12 * np.linalg.norm(want)


def test_c3_mass_total_and_mixed_error(c3):
    mesh = c3["mesh"]
    rho = 2.7e-9
    M = mesh.bilform(L.DOT, L.RULE_TET, 4, (rho * np.eye(3)).ravel())
    _, _, mz = M.csc_arrays(base=0)
    M.close()
    # u^T M u for a unit translation: rho times the rule's volume, sum over
    # elements and points of w det J (the basis functions sum to one), which
    # the oracle's load vector of a unit force also adds up
    total = mz.sum() / 3.0
    vol = O.linform_dot(O.T10, c3["conn"], c3["xyz"], 1, [1.0]).sum()
    say("C3 mass total / rho", total / rho, "rule volume", vol)
    assert abs(total / rho - vol) <= 1e-10 * vol
    Mm = mesh.bilform(L.DOT, L.RULE_TET, 4, (rho * np.eye(3)).ravel(), precision=L.MIXED)
    _, _, mm = Mm.csc_arrays(base=0)
    Mm.close()
    err = np.linalg.norm(mm - mz) / np.linalg.norm(mz)
    say("C3 mass mixed rel Frobenius", err)
    assert err <= 1e-6
