"""GPU tests of the argument checks that guard device kernels: a rejected DOF
numbering leaves the mesh as it was, nodes of no element keep whatever DOF
numbers they carry (the reference never reads them), the device sparse()
refuses out-of-range indices before any kernel could write out of bounds,
coordinates updated right after an asynchronous re-assembly are ordered after
it, and non-symmetric material matrices are refused.  Every case runs through
the C ABI; none makes a kernel fault (the checks are host-side or flag-only)."""
import ctypes
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
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb else 1.0)


def h8_mesh(n, amp, seed):
    conn, xyz = core.gen_h8block(1.0, 1.0, 1.0, *n)
    return conn, core.gen_perturb(xyz, [0, 0, 0], [1, 1, 1], amp / max(n), seed)


D200 = O.iso_D(200e3, 0.3)


def test_rejected_numbering_leaves_the_mesh_unchanged(ctx):
    say("rejected numbering")
    conn, xyz = h8_mesh((3, 3, 2), 0.1, 3)
    nn = xyz.shape[0]
    m = core.DeviceMesh(ctx, L.H8, conn, xyz)
    m.set_dofs(3)
    K0 = m.bilform(L.LIN_ELASTIC, L.RULE_GAUSS, 2, D200)
    ref = K0.csc_arrays(1)
    F0 = m.linform_dot(L.RULE_GAUSS, 2, [1.0, 2.0, -3.0], 3 * nn)
    for bad, text in ((3 * nn + 1, "Row degree of freedom > size"), (0, "Row degree of freedom < 1")):
        d = np.arange(1, 3 * nn + 1, dtype=np.int64).reshape(nn, 3)[::-1].copy()  # a permutation: COO path
        d[int(conn[1, 4]) - 1, 1] = bad
        with pytest.raises(L.FeasmError) as ei:
            m.set_dofs(3, d, 3 * nn)
        assert ei.value.code == L.EDOF and str(ei.value) == text
        # the mesh still holds the node-major numbering and its pattern
        K1 = m.bilform(L.LIN_ELASTIC, L.RULE_GAUSS, 2, D200)
        for a, b in zip(K1.csc_arrays(1), ref):
            assert np.array_equal(a, b)
        K1.close()
        assert np.array_equal(m.linform_dot(L.RULE_GAUSS, 2, [1.0, 2.0, -3.0], 3 * nn), F0)
    K0.close()
    m.close()


def test_nodes_of_no_element_keep_unused_dof_numbers(ctx):
    """A node that no element references may carry a DOF number outside
    1..nalldofs (assemble! never sees it): the numbering is accepted, the
    stiffness and the load vector match the oracle, and the load vector's
    gather does not write for that node."""
    say("node of no element")
    conn, xyz = h8_mesh((3, 2, 2), 0.1, 4)
    nn = xyz.shape[0]
    xyz2 = np.vstack([xyz, [[5.0, 5.0, 5.0]]])  # node nn+1, in no element
    rng = np.random.default_rng(8)
    perm = rng.permutation(3 * nn) + 1
    d = np.vstack([perm.reshape(nn, 3), [[10 ** 9, -7, 0]]]).astype(np.int64)
    m = core.DeviceMesh(ctx, L.H8, conn, xyz2)
    m.set_dofs(3, d, 3 * nn)
    K = m.bilform(L.LIN_ELASTIC, L.RULE_GAUSS, 2, D200)
    cp, rv, nz = K.csc_arrays(1)
    ocp, orv, onz = O.assemble_csc(O.H8, O.LIN_ELASTIC, conn, xyz2, 3, D200.ravel(), dofnums=d, nalldofs=3 * nn)
    assert np.array_equal(cp, ocp) and np.array_equal(rv, orv)
    assert rel_fro(nz, onz) <= 1e-12
    f = [0.5, -1.0, 2.0]
    F = m.linform_dot(L.RULE_GAUSS, 2, f, 3 * nn)
    OF = O.linform_dot(O.H8, conn, xyz2, 3, f, dofnums=d, nalldofs=3 * nn)
    assert rel_fro(F, OF) <= 1e-14
    K.close()
    m.close()


def test_sparse_device_refuses_out_of_range_indices(ctx, hiprt):
    """One bad index among valid triples: FEASM_EDOF before any output is
    written (the key pass only raises a flag); zero rows with triples too."""
    from test_gpu_device_entry_points import DeviceBuffers

    rng = np.random.default_rng(3)
    m_, n_, nt = 11, 7, 500
    I0 = rng.integers(0, m_, nt).astype(np.int32)
    J0 = rng.integers(0, n_, nt).astype(np.int32)
    V = rng.normal(size=nt)
    buf = DeviceBuffers(hiprt)
    try:
        for what, I, J, mm, nn in (("row = m", np.where(np.arange(nt) == 17, m_, I0), J0, m_, n_),
                                   ("negative column", I0, np.where(np.arange(nt) == 401, -1, J0), m_, n_),
                                   ("column = n", I0, np.where(np.arange(nt) == 0, n_, J0), m_, n_),
                                   ("m = 0", I0, J0, 0, n_)):
            say("sparse_device:", what)
            dI = buf.upload(np.ascontiguousarray(I, dtype=np.int32))
            dJ = buf.upload(np.ascontiguousarray(J, dtype=np.int32))
            dV = buf.upload(V)
            h = ctypes.c_void_p()
            rc = L.lib().feasm_sparse_device(ctx.h, nt, dI, dJ, dV, mm, nn, ctypes.byref(h))
            assert rc == L.EDOF, (what, rc)
            assert b"out of range" in L.lib().feasm_last_error(ctx.h)
        # the context still works afterwards
        dI, dJ, dV = buf.upload(I0), buf.upload(J0), buf.upload(V)
        h = ctypes.c_void_p()
        assert L.lib().feasm_sparse_device(ctx.h, nt, dI, dJ, dV, m_, n_, ctypes.byref(h)) == 0
        M = core.DeviceMatrix(ctx, h)
        ocp, orv, onz = O.sparse(I0.astype(np.int64) + 1, J0.astype(np.int64) + 1, V, m_, n_)
        cp, rv, nz = M.csc_arrays(1)
        assert np.array_equal(cp, ocp) and np.array_equal(rv, orv) and np.array_equal(nz, onz)
        M.close()
    finally:
        buf.free()


@pytest.mark.parametrize("path", ["staged", "fused"])
def test_set_xyz_right_after_an_asynchronous_reassembly(ctx, path):
    """bilform_into returns before its kernels finish; set_xyz issued
    immediately afterwards is ordered after them, so the first re-assembly
    sees the old coordinates and the second the new ones."""
    say("set_xyz ordering", path)
    conn, xyz = h8_mesh((8, 7, 6), 0.1, 1)
    _, xyz2 = h8_mesh((8, 7, 6), 0.1, 2)
    m = core.DeviceMesh(ctx, L.H8, conn, xyz)
    m.set_dofs(3)
    m.set_numeric_path(path)
    K = m.bilform(L.LIN_ELASTIC, L.RULE_GAUSS, 2, D200)
    K2 = m.bilform(L.LIN_ELASTIC, L.RULE_GAUSS, 2, D200)
    for _ in range(3):
        m.bilform_into(K, L.LIN_ELASTIC, L.RULE_GAUSS, 2, D200)
        m.set_xyz(xyz2)                      # no synchronisation in between
        m.bilform_into(K2, L.LIN_ELASTIC, L.RULE_GAUSS, 2, D200)
        ocp, orv, onz1 = O.assemble_csc(O.H8, O.LIN_ELASTIC, conn, xyz, 3, D200.ravel())
        _, _, onz2 = O.assemble_csc(O.H8, O.LIN_ELASTIC, conn, xyz2, 3, D200.ravel())
        assert rel_fro(K.csc_arrays(1)[2], onz1) <= 1e-12
        assert rel_fro(K2.csc_arrays(1)[2], onz2) <= 1e-12
        m.set_xyz(xyz)
    K.close()
    K2.close()
    m.close()


@pytest.mark.parametrize("form,ndn,coeff", [
    (L.LIN_ELASTIC, 3, D200 + np.triu(np.full((6, 6), 1.0), 1)),
    (L.DIFFUSION_MAT, 1, np.array([[2.0, 0.5, 0.0], [0.25, 2.0, 0.0], [0.0, 0.0, 1.0]])),
])
def test_non_symmetric_material_is_refused(ctx, form, ndn, coeff):
    """The reference forms the upper triangle of each element matrix and
    mirrors it; the node-block kernels need a symmetric D or kappa, so a
    non-symmetric one is refused rather than assembled differently."""
    conn, xyz = h8_mesh((2, 2, 2), 0.1, 5)
    m = core.DeviceMesh(ctx, L.H8, conn, xyz)
    m.set_dofs(ndn)
    for path in ("auto", "coo"):
        with pytest.raises(L.FeasmError) as ei:
            m.bilform(form, L.RULE_GAUSS, 2, coeff, path=path)
        assert ei.value.code == L.EUNSUPPORTED and "symmetric" in str(ei.value)
    m.close()
