"""The generic protocol (startassembly / assemble / makematrix) of the four
sparse-matrix assemblers and the free-free block wrapper, with makematrix on
the device, against the reference's own known answers:
  - the 7x7 testA through SysmatAssemblerSparse and SysmatAssemblerSparseSymm
    (test/test_basics.jl:78-133), the symmetric result bitwise the oracle's
    S + transpose(S);
  - the diagonal of testA through SysmatAssemblerSparseDiag (:135-178);
  - SysmatAssemblerSparseHRZLumpingSymm against the test's hand loop
    (:920-960), with a repeated DOF in one element;
  - SysmatAssemblerFFBlock's matrix_blocked_ff, square and not;
  - element matrices of a mesh (the oracle's) through the symmetric protocol,
    bitwise the oracle's symmetric assembly."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from feasm import _lib as L
from feasm import finetools as F
import oracle as O
from test_oracle_assembly import M1, M2

pytestmark = pytest.mark.gpu

TESTA = np.array([
    [2.85928, 1.21875, 0.891063, 0.891614, 2.56958, 0.0, 0.0],
    [1.21875, 1.15515, 0.716396, 0.0714644, 1.56825, 0.0, 0.0],
    [0.891063, 0.716396, 0.936979, 0.0, 1.36026, 0.0, 0.0],
    [0.891614, 0.0714644, 0.0, 0.661253, 0.813892, 0.0, 0.0],
    [2.56958, 1.56825, 1.36026, 0.813892, 4.15934, 0.0, 0.0],
    [0.0] * 7, [0.0] * 7])
I1, I2 = [5, 2, 1, 4], [2, 3, 1, 5]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def lower_coo(m, dofs):
    I, J, V = [], [], []
    for j in range(len(dofs)):
        for i in range(j, len(dofs)):
            I.append(dofs[i])
            J.append(dofs[j])
            V.append(m[i, j])
    return I, J, V


def test_kat_testA_sparse_and_symm(ctx):
    m1, m2 = M1.T @ M1, M2.T @ M2
    out = {}
    for name, cls in (("sparse", F.SysmatAssemblerSparse), ("symm", F.SysmatAssemblerSparseSymm)):
        a = cls(0.0, ctx=ctx)
        F.startassembly(a, 5, 5, 3, 7, 7)
        F.assemble(a, m1, I1, I1)
        F.assemble(a, m2, I2, I2)
        out[name] = F.makematrix(a)
        assert np.abs(TESTA - out[name].toarray()).max() < 1e-5, name
        a.close()
    A, Au = out["symm"].toarray(), out["sparse"].toarray()
    assert np.abs(A - Au).max() < 1e-5
    assert np.array_equal(A, A.T)
    # bitwise the oracle's restatement of S + transpose(S)
    a = lower_coo(m1, I1)
    b = lower_coo(m2, I2)
    ocp, orv, onz = O.symm_makematrix(a[0] + b[0], a[1] + b[1], a[2] + b[2], 7)
    S = out["symm"]
    assert np.array_equal(S.indptr, ocp - 1) and np.array_equal(S.indices, orv - 1) and np.array_equal(S.data, onz)


def test_symm_protocol_drops_zero_sums(ctx):
    # lower triples (2,1) = 1 and, from a second element in the other local
    # order, (1,2) = -1: they cancel in S + S' and are not stored
    a = F.SysmatAssemblerSparseSymm(0.0, ctx=ctx)
    F.startassembly(a, 2, 2, 2, 2, 2)
    F.assemble(a, np.array([[0.0, 0.0], [1.0, 3.0]]), [1, 2], [1, 2])
    F.assemble(a, np.array([[0.0, 0.0], [-1.0, 0.0]]), [2, 1], [2, 1])
    S = F.makematrix(a)
    assert S.indptr.tolist() == [0, 0, 1] and S.indices.tolist() == [1]
    assert S.toarray().tolist() == [[0.0, 0.0], [0.0, 3.0]]
    a.close()


def test_kat_diag_assembler(ctx):
    a = F.SysmatAssemblerSparseDiag(0.0, ctx=ctx)
    F.startassembly(a, 5, 5, 3, 7, 7)
    F.assemble(a, M1.T @ M1, I1, I1)
    F.assemble(a, M2.T @ M2, I2, I2)
    A = F.makematrix(a)
    assert np.abs(np.diag(np.diag(TESTA)) - A.toarray()).max() < 1e-5
    assert A.nnz == 5  # one entry per touched DOF, duplicates summed
    a.close()


def test_hrz_protocol_against_the_hand_loop(ctx):
    N, dim = 30, 8
    rng = np.random.default_rng(13)
    e = rng.random((dim, dim))
    e = e + e.T
    Mref = np.zeros((N, N))

    def add(d):
        em2 = e.sum(axis=0).sum()
        ffactor = em2 / np.trace(e)
        for i, di in enumerate(d):
            Mref[di - 1, di - 1] += ffactor * e[i, i]

    a = F.SysmatAssemblerSparseHRZLumpingSymm(0.0, ctx=ctx)
    F.startassembly(a, dim, dim, 3, N, N)
    for d in ([10, 29, 15, 1, 7, 3, 6, 2], [15, 1, 7, 11, 29, 3, 8, 4], [25, 1, 17, 13, 30, 13, 18, 14]):
        F.assemble(a, e, d, d)
        add(d)
    M = F.makematrix(a).toarray()
    s = np.abs(np.diag(M) - np.diag(Mref)).sum()
    assert s / np.linalg.norm(e) <= 1e-9
    assert np.count_nonzero(M - np.diag(np.diag(M))) == 0
    a.close()


@pytest.mark.parametrize("r,c", [(4, 4), (5, 3), (0, 3), (7, 7)])
def test_ffblock_protocol(ctx, r, c):
    m1, m2 = M1.T @ M1, M2.T @ M2
    a = F.SysmatAssemblerFFBlock(r, c, ctx=ctx)
    F.startassembly(a, 4, 4, 2, 7, 7)
    F.assemble(a, m1, I1, I1)
    F.assemble(a, m2, I2, I2)
    B = F.makematrix(a)
    # the plain assembly's leading block, stored entries kept
    p = F.SysmatAssemblerSparse(0.0, ctx=ctx)
    F.startassembly(p, 4, 4, 2, 7, 7)
    F.assemble(p, m1, I1, I1)
    F.assemble(p, m2, I2, I2)
    full = F.makematrix(p)
    want = full[:r, :c] if r and c else sp.csc_matrix((r, c))
    assert B.shape == (r, c)
    assert np.array_equal(B.toarray(), want.toarray())
    assert B.nnz == want.nnz


def test_ffblock_too_many_rows(ctx):
    a = F.SysmatAssemblerFFBlock(8, ctx=ctx)
    F.startassembly(a, 4, 4, 1, 7, 7)
    F.assemble(a, M1.T @ M1, I1, I1)
    with pytest.raises(L.FeasmError) as ei:
        F.makematrix(a)
    assert str(ei.value) == "The ff block has too many rows"


@pytest.mark.parametrize("etype,form,ndn,coeff", [
    (O.H8, O.LIN_ELASTIC, 3, O.iso_D(200e3, 0.3)),
    (O.T10, O.DIFFUSION_ISO, 1, np.array([1.3])),
    (O.H8, O.DOT, 3, 7.8e-9 * np.eye(3)),
])
def test_symm_protocol_with_element_matrices_matches_oracle(ctx, etype, form, ndn, coeff):
    """What an application's own element loop does with the default
    assembler: every element matrix (the oracle's) through startassembly /
    assemble / makematrix of SysmatAssemblerSparseSymm."""
    cells = (3, 2, 2) if etype == O.H8 else (2, 2, 1)
    conn, xyz = O.block_mesh(etype, cells, L=1.0, amp_frac=0.1, seed=6)
    nn, npe = xyz.shape[0], conn.shape[1]
    a = F.SysmatAssemblerSparseSymm(0.0, ctx=ctx)
    F.startassembly(a, npe * ndn, npe * ndn, conn.shape[0], nn * ndn, nn * ndn)
    for e in range(conn.shape[0]):
        X = xyz[conn[e] - 1]
        Ke = O.elmat(etype, form, ndn, X, np.ravel(coeff))
        d = ((conn[e] - 1)[:, None] * ndn + np.arange(ndn)[None, :] + 1).ravel()
        F.assemble(a, Ke, d, d)
    S = F.makematrix(a)
    ocp, orv, onz = O.assemble_csc_symm(etype, form, conn, xyz, ndn, np.ravel(coeff))
    assert np.array_equal(S.indptr, ocp - 1) and np.array_equal(S.indices, orv - 1)
    assert np.array_equal(S.data, onz)
    a.close()
