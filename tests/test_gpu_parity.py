"""GPU parity: the HIP assembly against the oracle on the same seeded meshes.

Bar: colptr/rowval bit-exact; FP64 nzval within 1e-12 relative Frobenius;
mixed precision within 1e-6.  Every case runs through the C ABI.
"""
import ctypes
import sys

import numpy as np
import pytest

import feasm
from feasm import _lib as L
from feasm import core
import oracle as O

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 1e-6


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_fro(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb else 1.0)


def mesh_for(etype, n, amp, seed):
    if etype == O.H8:
        conn, xyz = core.gen_h8block(1.0, 1.0, 1.0, *n)
        if amp:
            xyz = core.gen_perturb(xyz, [0, 0, 0], [1, 1, 1], amp / max(n), seed)
    else:
        h = 1.0 / max(n)
        conn, xyz = core.gen_t10block(1.0, 1.0, 1.0, *n, amp * h, 0.3 * amp * h, seed)
    return conn, xyz


def rule_for(etype):
    return (L.RULE_GAUSS, 2) if etype == O.H8 else (L.RULE_TET, 4)


D200 = O.iso_D(200e3, 0.3)
D70 = O.iso_D(70e3, 0.33)
_rng = np.random.default_rng(5)
_A = _rng.normal(size=(3, 3))
KAPPA = _A @ _A.T + np.eye(3)
_B = _rng.normal(size=(6, 6))
DANISO = _B @ _B.T + 6 * np.eye(6)

CASES = [
    # name, etype, cells, amp, form, ndn, coeff
    ("C1 H8 diffusion 10^3", O.H8, (10, 10, 10), 0.1, L.DIFFUSION_ISO, 1, [1.0]),
    ("H8 elasticity", O.H8, (6, 5, 4), 0.15, L.LIN_ELASTIC, 3, D200),
    ("H8 anisotropic elasticity", O.H8, (3, 4, 3), 0.15, L.LIN_ELASTIC, 3, DANISO),
    ("H8 matrix diffusion", O.H8, (5, 4, 3), 0.1, L.DIFFUSION_MAT, 1, KAPPA),
    ("H8 mass 3", O.H8, (4, 4, 4), 0.1, L.DOT, 3, 7.8e-9 * np.eye(3)),
    ("H8 dot 1", O.H8, (4, 3, 5), 0.1, L.DOT, 1, [2.5]),
    ("T10 elasticity", O.T10, (3, 3, 2), 0.1, L.LIN_ELASTIC, 3, D70),
    ("T10 mass 3", O.T10, (3, 2, 3), 0.1, L.DOT, 3, 2.7e-9 * np.eye(3)),
    ("T10 diffusion", O.T10, (3, 3, 3), 0.1, L.DIFFUSION_ISO, 1, [1.7]),
    ("H8 single element", O.H8, (1, 1, 1), 0.0, L.LIN_ELASTIC, 3, D200),
]


def oracle_csc(etype, form, conn, xyz, ndn, coeff):
    return O.assemble_csc(etype, form, conn, xyz, ndn, np.ravel(coeff))


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("path", ["fused", "staged", "coo"])
def test_fp64_parity(ctx, case, path):
    name, etype, n, amp, form, ndn, coeff = case
    say("case", name, path)
    conn, xyz = mesh_for(etype, n, amp, 1234)
    ocp, orv, onz = oracle_csc(etype, form, conn, xyz, ndn, coeff)
    mesh = core.DeviceMesh(ctx, etype, conn, xyz)
    mesh.set_dofs(ndn)
    if path != "coo":
        mesh.set_numeric_path(path)
    rule, order = rule_for(etype)
    M = mesh.bilform(form, rule, order, np.ravel(coeff), precision=L.FP64, path=path)
    cp, rv, nz = M.csc_arrays(base=1)
    assert np.array_equal(cp, ocp)
    assert np.array_equal(rv, orv)
    err = rel_fro(nz, onz)
    say("  rel Frobenius", err)
    assert err <= FP64_TOL
    if form != L.DOT or np.allclose(np.reshape(coeff, (ndn, ndn)), np.reshape(coeff, (ndn, ndn)).T):
        import scipy.sparse as sp
        K = sp.csc_matrix((nz, rv - 1, cp - 1), shape=(len(cp) - 1,) * 2)
        assert (K - K.T).count_nonzero() == 0  # bit-exact symmetry (test/test_forms.jl:442)


@pytest.mark.parametrize("case", [c for c in CASES if c[1] == O.T10 or c[4] == L.LIN_ELASTIC],
                         ids=lambda c: c[0])
@pytest.mark.parametrize("path", ["fused", "staged"])
def test_mixed_precision_error(ctx, case, path):
    name, etype, n, amp, form, ndn, coeff = case
    say("mixed", name, path)
    conn, xyz = mesh_for(etype, n, amp, 99)
    ocp, orv, onz = oracle_csc(etype, form, conn, xyz, ndn, coeff)
    mesh = core.DeviceMesh(ctx, etype, conn, xyz)
    mesh.set_dofs(ndn)
    mesh.set_numeric_path(path)
    rule, order = rule_for(etype)
    M = mesh.bilform(form, rule, order, np.ravel(coeff), precision=L.MIXED)
    cp, rv, nz = M.csc_arrays(base=1)
    assert np.array_equal(cp, ocp) and np.array_equal(rv, orv)
    err = rel_fro(nz, onz)
    say("  mixed rel Frobenius", err)
    assert err <= MIXED_TOL


def test_reassembly_reuses_pattern_and_is_deterministic(ctx):
    conn, xyz = mesh_for(O.H8, (5, 5, 5), 0.15, 7)
    mesh = core.DeviceMesh(ctx, O.H8, conn, xyz)
    mesh.set_dofs(3)
    M = mesh.bilform(L.LIN_ELASTIC, L.RULE_GAUSS, 2, D200)
    a = M.csc_arrays(1)
    mesh.bilform_into(M, L.LIN_ELASTIC, L.RULE_GAUSS, 2, D70)
    b = M.csc_arrays(1)
    ocp, orv, onz = oracle_csc(O.H8, L.LIN_ELASTIC, conn, xyz, 3, D70)
    assert np.array_equal(b[0], ocp) and np.array_equal(b[1], orv)
    assert rel_fro(b[2], onz) <= FP64_TOL
    mesh.bilform_into(M, L.LIN_ELASTIC, L.RULE_GAUSS, 2, D200)
    c = M.csc_arrays(1)
    assert np.array_equal(a[2], c[2])  # bitwise repeatable (test/test_basics.jl:3016-3051)


def test_general_dof_numbering_goes_through_coo(ctx):
    conn, xyz = mesh_for(O.H8, (3, 3, 3), 0.1, 3)
    nn = xyz.shape[0]
    rng = np.random.default_rng(0)
    perm = rng.permutation(nn)
    dof = np.zeros((nn, 3), np.int64)
    nxt = 1
    for i in perm:  # numberdofs!(u, entperm)
        dof[i] = [nxt, nxt + 1, nxt + 2]
        nxt += 3
    I, J, V = O.bilform_coo(O.H8, O.LIN_ELASTIC, conn, xyz, 3, dof, 3 * nn, D200.ravel())
    ocp, orv, onz = O.sparse(I, J, V, 3 * nn, 3 * nn)
    mesh = core.DeviceMesh(ctx, O.H8, conn, xyz)
    mesh.set_dofs(3, dof, 3 * nn)
    M = mesh.bilform(L.LIN_ELASTIC, L.RULE_GAUSS, 2, D200)
    cp, rv, nz = M.csc_arrays(1)
    assert np.array_equal(cp, ocp) and np.array_equal(rv, orv)
    assert rel_fro(nz, onz) <= FP64_TOL


def test_sparse_device_order_and_zeros(ctx):
    # left fold in input order: (1e16 + 1) - 1e16 == 0; explicit zeros kept
    M = core.sparse(ctx, np.array([2, 2, 2, 1]), np.array([1, 1, 1, 3]), np.array([1e16, 1.0, -1e16, 0.0]), 3, 3)
    cp, rv, nz = M.csc_arrays(1)
    assert cp.tolist() == [1, 2, 2, 3] and rv.tolist() == [2, 1]
    assert nz[0] == 0.0 and nz[1] == 0.0
    rng = np.random.default_rng(2)
    m, n, t = 101, 77, 50000
    I = rng.integers(1, m + 1, t)
    J = rng.integers(1, n + 1, t)
    V = rng.normal(size=t)
    ocp, orv, onz = O.sparse(I, J, V, m, n)
    M = core.sparse(ctx, I, J, V, m, n)
    cp, rv, nz = M.csc_arrays(1)
    assert np.array_equal(cp, ocp) and np.array_equal(rv, orv)
    assert np.array_equal(nz, onz)  # same additions in the same order: bitwise


def test_finetools_mirror_protocol_kats(ctx):
    from feasm.finetools import SysmatAssemblerSparse, assemble, makematrix, setnomatrixresult, startassembly

    m1 = np.array([[0.24406, 0.599773, 0.833404, 0.0420141],
                   [0.786024, 0.00206713, 0.995379, 0.780298],
                   [0.845816, 0.198459, 0.355149, 0.224996]])
    m2 = np.array([[0.146618, 0.53471, 0.614342, 0.737833],
                   [0.479719, 0.41354, 0.00760941, 0.836455],
                   [0.254868, 0.476189, 0.460794, 0.00919633],
                   [0.159064, 0.261821, 0.317078, 0.77646],
                   [0.643538, 0.429817, 0.59788, 0.958909]])
    # test/test_basics.jl:1577-1613 (massnomat1)
    a = SysmatAssemblerSparse(0.0, True, ctx=ctx)
    startassembly(a, 5, 5, 3, 7, 7)
    assemble(a, m1, [1, 7, 5], [5, 2, 1, 4])
    assemble(a, m2, [2, 3, 1, 7, 5], [6, 7, 3, 4])
    A = makematrix(a)
    assert A.nnz == 0 and A.shape == (7, 7)
    setnomatrixresult(a, False)
    A = makematrix(a).toarray()
    for (i, j, v) in ((1, 1, 0.833404), (5, 1, 0.355149), (7, 6, 0.159064), (3, 7, 0.41354), (7, 7, 0.261821)):
        assert abs(A[i - 1, j - 1] - v) < 1e-12
    # test/test_basics.jl:78-133 (mmassembly2ya1)
    testA = np.array([
        [2.85928, 1.21875, 0.891063, 0.891614, 2.56958, 0.0, 0.0],
        [1.21875, 1.15515, 0.716396, 0.0714644, 1.56825, 0.0, 0.0],
        [0.891063, 0.716396, 0.936979, 0.0, 1.36026, 0.0, 0.0],
        [0.891614, 0.0714644, 0.0, 0.661253, 0.813892, 0.0, 0.0],
        [2.56958, 1.56825, 1.36026, 0.813892, 4.15934, 0.0, 0.0],
        [0.0] * 7, [0.0] * 7])
    a = SysmatAssemblerSparse(0.0, ctx=ctx)
    startassembly(a, 5, 5, 3, 7, 7)
    assemble(a, m1.T @ m1, [5, 2, 1, 4], [5, 2, 1, 4])
    assemble(a, m2.T @ m2, [2, 3, 1, 5], [2, 3, 1, 5])
    assert np.abs(makematrix(a).toarray() - testA).max() < 1e-5


def test_finetools_mirror_bilforms(ctx):
    from feasm.finetools import (DataCache, DeforModelRed3D, FEMMBase, GaussRule, H8block, IntegDomain, NodalField,
                                 SysmatAssemblerSparse, bilform_diffusion, bilform_dot, bilform_lin_elastic,
                                 numberdofs)

    # test/test_forms.jl:331-367 patch test and test/test_basics.jl:962-1020 Gram
    W, Lx, t = 11.1, 12.0, 7.32
    fens, fes = H8block(Lx, W, t, 2, 4, 3)
    geom = NodalField(fens.xyz)
    femm = FEMMBase(IntegDomain(fes, GaussRule(3, 2)), ctx=ctx)
    a_, b_, c_, d_ = (-0.33, 2 / 3, -1.67, 2 / 7)
    X, Y, Z = fens.xyz.T
    u = NodalField(np.column_stack([a_ + b_ * X + c_ * Y + d_ * Z, b_ + c_ * X + d_ * Y + a_ * Z,
                                    c_ + d_ * X + a_ * Y + b_ * Z]))
    numberdofs(u)
    mu = 0.13377
    G = bilform_lin_elastic(femm, SysmatAssemblerSparse(0.0, ctx=ctx), geom, u, DeforModelRed3D,
                            DataCache(np.diag([2 * mu] * 3 + [mu] * 3)))
    gradu = np.array([[b_, c_, d_], [c_, d_, a_], [d_, a_, b_]])
    sym = (gradu + gradu.T) / 2
    exact = 2 * mu * (W * Lx * t) * (sym ** 2).sum()
    v = u.values.ravel()
    assert abs(v @ (G @ v) - exact) / exact <= 1e-5
    assert (G - G.T).count_nonzero() == 0
    psi = NodalField(np.ones((fens.count(), 1)))
    numberdofs(psi)
    Gm = bilform_dot(femm, SysmatAssemblerSparse(0.0, ctx=ctx), geom, psi, DataCache(np.eye(1)))
    one = np.ones(fens.count())
    assert abs(one @ (Gm @ one) - W * Lx * t) / (W * Lx * t) <= 1e-5
    K = bilform_diffusion(femm, SysmatAssemblerSparse(0.0, ctx=ctx), geom, psi, DataCache(1.0))
    assert abs(one @ (K @ one)) <= 1e-10


def test_errors_surface_with_reference_text(ctx):
    conn, xyz = mesh_for(O.H8, (2, 2, 2), 0.0, 0)
    mesh = core.DeviceMesh(ctx, O.H8, conn, xyz)
    nn = xyz.shape[0]
    dof = np.arange(1, nn + 1).reshape(nn, 1)
    dof[4, 0] = 0
    with pytest.raises(feasm.FeasmError, match="degree of freedom < 1"):
        mesh.set_dofs(1, dof, nn)
    mesh.set_dofs(1)
    with pytest.raises(feasm.FeasmError, match="Wrong dimensions"):
        mesh.bilform(L.LIN_ELASTIC, L.RULE_GAUSS, 2, D200)
    with pytest.raises(feasm.FeasmError, match="Gauss"):
        mesh.bilform(L.DIFFUSION_ISO, L.RULE_TET, 4, [1.0])
    bad = np.array([[1, 2, 3, 4, 5, 6, 7, nn + 5]], np.int64)
    with pytest.raises(feasm.FeasmError, match="missing node"):
        core.DeviceMesh(ctx, O.H8, bad, xyz)


@pytest.mark.parametrize("etype,n,form,ndn,coeff", [
    (O.H8, (9, 7, 8), L.LIN_ELASTIC, 3, D200),
    (O.T10, (5, 4, 5), L.LIN_ELASTIC, 3, D70),
    (O.H8, (12, 11, 10), L.DIFFUSION_ISO, 1, [1.3]),
    (O.T10, (4, 5, 4), L.DOT, 3, 2.7e-9 * np.eye(3)),
])
def test_fused_matches_staged_and_oracle_on_many_tiles(ctx, etype, n, form, ndn, coeff):
    # meshes large enough for many tiles with halos crossing tile borders
    say("fused vs staged", etype, n, form)
    conn, xyz = mesh_for(etype, n, 0.15 if etype == O.H8 else 0.1, 77)
    mesh = core.DeviceMesh(ctx, etype, conn, xyz)
    mesh.set_dofs(ndn)
    rule, order = rule_for(etype)
    mesh.set_numeric_path("fused")
    F = mesh.bilform(form, rule, order, np.ravel(coeff))
    ntiles, nhalo = mesh.tile_info()
    say("  tiles", ntiles, "halo elements", nhalo, "elements", conn.shape[0])
    assert ntiles > 4
    mesh.set_numeric_path("staged")
    S = mesh.bilform(form, rule, order, np.ravel(coeff))
    f, s = F.csc_arrays(1), S.csc_arrays(1)
    assert np.array_equal(f[0], s[0]) and np.array_equal(f[1], s[1])
    ocp, orv, onz = oracle_csc(etype, form, conn, xyz, ndn, coeff)
    assert np.array_equal(f[0], ocp) and np.array_equal(f[1], orv)
    assert rel_fro(f[2], onz) <= FP64_TOL and rel_fro(s[2], onz) <= FP64_TOL
    import scipy.sparse as sp
    K = sp.csc_matrix((f[2], f[1] - 1, f[0] - 1), shape=(len(f[0]) - 1,) * 2)
    assert (K - K.T).count_nonzero() == 0


def test_batched_launch_matches_single_and_oracle_with_graph_replays(ctx):
    """Each mesh's result is the same alone and in the batch (bitwise, same
    tiles), and every replay of the captured graph gives the eager result."""
    rng = np.random.default_rng(11)
    specs = []
    for i in range(6):
        n = tuple(int(x) for x in rng.integers(2, 7, 3))
        if i % 2 == 0:
            specs.append((n, L.DIFFUSION_ISO, 1, [float(rng.uniform(0.5, 2.0))]))
        else:
            specs.append((n, L.LIN_ELASTIC, 3, O.iso_D(float(rng.uniform(1e3, 1e5)), float(rng.uniform(0.2, 0.45)))))
    for precision in (L.FP64, L.MIXED):
        say("batch precision", precision)
        batch = core.Batch(ctx)
        meshes, singles = [], []
        for j, (n, form, ndn, coeff) in enumerate(specs):
            conn, xyz = mesh_for(O.H8, n, 0.1, 100 + j)
            mesh = core.DeviceMesh(ctx, O.H8, conn, xyz)
            mesh.set_dofs(ndn)
            singles.append(mesh.bilform(form, L.RULE_GAUSS, 2, np.ravel(coeff), precision=precision).csc_arrays(1))
            batch.add(mesh, form, L.RULE_GAUSS, 2, np.ravel(coeff), precision)
            meshes.append((mesh, conn, xyz, form, ndn, coeff))
        batch.prepare(use_graph=True)
        for rep in range(3):
            batch.run()
            ctx.synchronize()
            for j, (mesh, conn, xyz, form, ndn, coeff) in enumerate(meshes):
                got = batch.matrix(j).csc_arrays(1)
                assert np.array_equal(got[0], singles[j][0]) and np.array_equal(got[1], singles[j][1])
                assert np.array_equal(got[2], singles[j][2]), (rep, j)
            # eager work between replays (graph replays must not depend on it)
            meshes[0][0].bilform(specs[0][1], L.RULE_GAUSS, 2, np.ravel(specs[0][3]))
        for j, (mesh, conn, xyz, form, ndn, coeff) in enumerate(meshes):
            ocp, orv, onz = oracle_csc(O.H8, form, conn, xyz, ndn, coeff)
            err = rel_fro(batch.matrix(j).csc_arrays(1)[2], onz)
            assert err <= (FP64_TOL if precision == L.FP64 else MIXED_TOL), (j, err)
        st = batch.stats()
        assert st["meshes"] == 6 and st["tiles"] > 0
        batch.close()


def test_server_results_same_alone_and_in_the_mix():
    """Continuous batching: each request's matrix is bitwise the same when it
    runs alone and when it runs mixed with others (and matches the oracle)."""
    rng = np.random.default_rng(21)
    reqs = []
    for i in range(8):
        if i % 2 == 0:
            n = tuple(int(x) for x in rng.integers(3, 9, 3))
            conn, xyz = mesh_for(O.H8, n, 0.1, 300 + i)
            et, rule = O.H8, (L.RULE_GAUSS, 2)
        else:
            n = tuple(int(x) for x in rng.integers(2, 5, 3))
            conn, xyz = mesh_for(O.T10, n, 0.1, 300 + i)
            et, rule = O.T10, (L.RULE_TET, 4)
        if i % 4 < 2:
            form, ndn, coeff = L.DIFFUSION_ISO, 1, [float(rng.uniform(0.5, 2))]
        else:
            form, ndn, coeff = L.LIN_ELASTIC, 3, O.iso_D(float(rng.uniform(1e3, 1e5)), float(rng.uniform(0.2, 0.45)))
        reqs.append((et, conn, xyz, ndn, form, rule, np.ravel(coeff)))
    for precision, tol in ((L.FP64, FP64_TOL), (L.MIXED, MIXED_TOL)):
        say("server precision", precision)
        alone = []
        srv = core.Server(0, workers=1, hbm_fraction=0.5)
        for (et, conn, xyz, ndn, form, rule, coeff) in reqs:
            rid, _ = srv.submit_mesh(et, conn, xyz, ndn, form, rule[0], rule[1], coeff, precision)
            st, lat, _ = srv.wait(rid)
            assert st == 2
            alone.append(srv.result(rid).csc_arrays(1))
            srv.release(rid)
        srv.close()
        srv = core.Server(0, workers=4, hbm_fraction=0.5)
        ids = []
        for k in range(3):  # each request three times, interleaved, some kept resident
            for j, (et, conn, xyz, ndn, form, rule, coeff) in enumerate(reqs):
                rid, mid = srv.submit_mesh(et, conn, xyz, ndn, form, rule[0], rule[1], coeff, precision,
                                           keep=(k == 0))
                ids.append((j, rid, mid if k == 0 else None))
        mids = {}
        for j, rid, mid in ids:
            st, lat, _ = srv.wait(rid)
            assert st == 2, (j, st)
            got = srv.result(rid).csc_arrays(1)
            for x, y in zip(got, alone[j]):
                assert np.array_equal(x, y), j
            srv.release(rid)
            if mid is not None:
                mids[j] = mid
        # re-assembly of resident meshes with the original coefficients
        re = [(j, srv.submit_reassembly(mids[j], reqs[j][4], reqs[j][5][0], reqs[j][5][1], reqs[j][6], precision))
              for j in mids]
        for j, rid in re:
            st, lat, _ = srv.wait(rid)
            assert st == 2
            got = srv.result(rid).csc_arrays(1)
            assert np.array_equal(got[2], alone[j][2])
            srv.release(rid)
        st = srv.stats()
        assert st["completed"] == len(ids) + len(re) and st["residents"] == len(mids)
        srv.close()
        # and against the oracle
        for j, (et, conn, xyz, ndn, form, rule, coeff) in enumerate(reqs):
            ocp, orv, onz = oracle_csc(et, form, conn, xyz, ndn, coeff)
            assert np.array_equal(alone[j][0], ocp) and np.array_equal(alone[j][1], orv)
            assert rel_fro(alone[j][2], onz) <= tol


@pytest.mark.parametrize("etype", [O.H8, O.T10])
def test_moved_nodes_reassemble_with_the_new_coordinates(ctx, etype):
    """set_xyz after a fused assembly: the tiles' copies of the coordinates
    follow, and re-assembly into the same pattern matches the oracle on the
    moved mesh (fused and staged paths)."""
    n = (4, 3, 3) if etype == O.H8 else (3, 2, 2)
    conn, xyz = mesh_for(etype, n, 0.1, 1)
    _, xyz2 = mesh_for(etype, n, 0.1, 2)  # same topology, other perturbation
    rule = rule_for(etype)
    D = D200 if etype == O.H8 else D70
    for path in ("fused", "staged"):
        say("moved nodes", etype, "path", path)
        m = core.DeviceMesh(ctx, etype, conn, xyz)
        m.set_dofs(3)
        m.set_numeric_path(path)
        K = m.bilform(L.LIN_ELASTIC, rule[0], rule[1], D)
        m.set_xyz(xyz2)
        m.bilform_into(K, L.LIN_ELASTIC, rule[0], rule[1], D)
        cp, rv, nz = K.csc_arrays(1)
        ocp, orv, onz = oracle_csc(etype, L.LIN_ELASTIC, conn, xyz2, 3, D)
        assert np.array_equal(cp, ocp) and np.array_equal(rv, orv)
        assert rel_fro(nz, onz) <= FP64_TOL
        K.close()
        m.close()


@pytest.mark.parametrize("etype,n,form,ndn,coeff", [
    (O.H8, (7, 6, 5), L.LIN_ELASTIC, 3, D200),
    (O.H8, (5, 6, 4), L.LIN_ELASTIC, 3, DANISO),
    (O.T10, (3, 3, 3), L.DOT, 3, 2.7e-9 * np.eye(3)),
    (O.T10, (3, 2, 3), L.DIFFUSION_ISO, 1, [1.7]),
])
def test_tile_parameters_do_not_change_the_values(ctx, etype, n, form, ndn, coeff):
    """Every block adds the same element matrices in ascending element order
    whatever the tiling, so the values are bitwise the same for every tile
    size, LDS budget and workgroup size (small workgroups run several block
    tasks per thread), and match the oracle."""
    conn, xyz = mesh_for(etype, n, 0.15 if etype == O.H8 else 0.1, 11)
    rule, order = rule_for(etype)
    ocp, orv, onz = oracle_csc(etype, form, conn, xyz, ndn, coeff)
    for prec, tol in ((L.FP64, FP64_TOL), (L.MIXED, MIXED_TOL)):
        ref = None
        for params in ((0, 0, 0), (8, 48, 256), (24, 96, 512), (255, 160, 1024)):
            say("tile params", etype, form, prec, params)
            m = core.DeviceMesh(ctx, etype, conn, xyz)
            m.set_dofs(ndn)
            m.set_numeric_path("fused")
            m.set_tile_params(*params)
            K = m.bilform(form, rule, order, np.ravel(coeff), precision=prec)
            cp, rv, nz = K.csc_arrays(1)
            assert np.array_equal(cp, ocp) and np.array_equal(rv, orv)
            assert rel_fro(nz, onz) <= tol
            if ref is None:
                ref = nz
            else:
                assert np.array_equal(nz, ref), params
            K.close()
            m.close()


@pytest.mark.parametrize("etype", [O.H8, O.T10])
def test_inverted_elements_match_the_oracle(ctx, etype):
    """Elements with J w < 0 contribute with that sign, as the reference's
    `Jac * w[j]` does (src/FEMMBaseModule.jl:1494-1528, 1800-1806: no check
    on the sign).  Every fifth element is mirrored (H8: bottom and top faces
    swapped; T10: vertices 1 and 2 and their edge nodes swapped), which sends
    the fused kernel through its signed path."""
    n = (5, 4, 3) if etype == O.H8 else (3, 3, 2)
    conn, xyz = mesh_for(etype, n, 0.1, 3)
    perm = [4, 5, 6, 7, 0, 1, 2, 3] if etype == O.H8 else [0, 2, 1, 3, 6, 5, 4, 7, 9, 8]
    conn = conn.copy()
    conn[::5] = conn[::5][:, perm]
    rule, order = rule_for(etype)
    D = D200 if etype == O.H8 else D70
    for form, ndn, coeff in ((L.LIN_ELASTIC, 3, D), (L.DIFFUSION_ISO, 1, [1.3]), (L.DOT, 3, 2.7e-9 * np.eye(3))):
        ocp, orv, onz = oracle_csc(etype, form, conn, xyz, ndn, coeff)
        for path in ("fused", "staged", "coo"):
            for prec, tol in ((L.FP64, FP64_TOL), (L.MIXED, MIXED_TOL)):
                if path == "coo" and prec == L.MIXED:
                    continue
                say("inverted", etype, form, path, prec)
                m = core.DeviceMesh(ctx, etype, conn, xyz)
                m.set_dofs(ndn)
                if path != "coo":
                    m.set_numeric_path(path)
                K = m.bilform(form, rule, order, np.ravel(coeff), precision=prec,
                              path="coo" if path == "coo" else "auto")
                cp, rv, nz = K.csc_arrays(1)
                assert np.array_equal(cp, ocp) and np.array_equal(rv, orv)
                assert rel_fro(nz, onz) <= tol, (form, path, prec, rel_fro(nz, onz))
                K.close()
                m.close()


def test_device_resident_mesh_matches_host_mesh_and_checks_connectivity(ctx, hiprt):
    """feasm_mesh_create_device: a mesh made from device arrays assembles
    bitwise like the same mesh made from host arrays, and a connectivity entry
    outside [0, nnodes) is rejected with the host path's error."""
    hip = hiprt
    conn, xyz = mesh_for(O.H8, (4, 3, 3), 0.1, 8)
    conn0 = np.ascontiguousarray(conn - 1, dtype=np.int32)
    xyzr = np.ascontiguousarray(xyz, dtype=np.float64)  # row-major nnodes x 3
    bufs = []

    def upload(a):
        p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        bufs.append(p)
        assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # hipMemcpyHostToDevice
        return p

    try:
        d_conn, d_xyz = upload(conn0), upload(xyzr)
        lib = L.lib()
        h = ctypes.c_void_p()
        rc = lib.feasm_mesh_create_device(ctx.h, L.H8, conn.shape[0], 8, d_conn, xyz.shape[0], d_xyz,
                                          ctypes.byref(h))
        assert rc == 0, lib.feasm_last_error(ctx.h)
        dev = core.DeviceMesh.__new__(core.DeviceMesh)  # wraps the handle made above
        dev.ctx, dev.h, dev.etype, dev.ndn = ctx, h, O.H8, 0
        dev.nelem, dev.npe, dev.nnodes = conn.shape[0], 8, xyz.shape[0]
        host = core.DeviceMesh(ctx, O.H8, conn, xyz)
        vals = []
        for m in (dev, host):
            m.set_dofs(3)
            K = m.bilform(L.LIN_ELASTIC, L.RULE_GAUSS, 2, D200)
            vals.append(K.csc_arrays(1))
            K.close()
        for a, b in zip(*vals):
            assert np.array_equal(a, b)
        host.close()
        dev.close()
        bad = conn0.copy()
        bad[3, 5] = xyz.shape[0]  # one past the last node
        d_bad = upload(bad)
        h2 = ctypes.c_void_p()
        rc = lib.feasm_mesh_create_device(ctx.h, L.H8, conn.shape[0], 8, d_bad, xyz.shape[0], d_xyz,
                                          ctypes.byref(h2))
        assert rc == L.EINVAL and not h2.value
    finally:
        for p in bufs:
            hip.hipFree(p)
