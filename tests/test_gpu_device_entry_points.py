"""GPU tests of the entry points that take or fill device arrays, and of the
batched COO assemble: each against its host-array counterpart, bitwise, and
against the oracle."""
import ctypes

import numpy as np
import pytest

from feasm import _lib as L
from feasm import core
import oracle as O

pytestmark = pytest.mark.gpu


class DeviceBuffers:
    def __init__(self, hip):
        self.hip, self.bufs = hip, []

    def upload(self, a):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        if a.nbytes:
            assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, out):
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def test_sparse_device_matches_host_sparse_and_oracle(ctx, hiprt):
    """feasm_sparse_device on 0-based int32 device triples = feasm_sparse on
    1-based host triples = the oracle's sparse(): duplicates summed left to
    right, explicit zeros kept."""
    rng = np.random.default_rng(21)
    m, n, nt = 37, 29, 4000
    I1 = rng.integers(1, m + 1, nt)
    J1 = rng.integers(1, n + 1, nt)
    V = rng.normal(size=nt)
    V[::17] = 0.0
    V[5] = 1e16
    V[6] = 1.0
    V[7] = -1e16
    I1[5:8] = 3
    J1[5:8] = 4
    host = core.sparse(ctx, I1, J1, V, m, n).csc_arrays(1)
    buf = DeviceBuffers(hiprt)
    try:
        dI = buf.upload(np.ascontiguousarray(I1 - 1, dtype=np.int32))
        dJ = buf.upload(np.ascontiguousarray(J1 - 1, dtype=np.int32))
        dV = buf.upload(np.ascontiguousarray(V, dtype=np.float64))
        h = ctypes.c_void_p()
        rc = L.lib().feasm_sparse_device(ctx.h, nt, dI, dJ, dV, m, n, ctypes.byref(h))
        assert rc == 0, L.lib().feasm_last_error(ctx.h)
        M = core.DeviceMatrix(ctx, h)
        dev = M.csc_arrays(1)
        M.close()
    finally:
        buf.free()
    for a, b in zip(host, dev):
        assert np.array_equal(a, b)
    ocp, orv, onz = O.sparse(I1, J1, V, m, n)
    assert np.array_equal(dev[0], ocp) and np.array_equal(dev[1], orv) and np.array_equal(dev[2], onz)


@pytest.mark.parametrize("etype", [O.H8, O.T10])
def test_linform_dot_device_matches_host_vector(ctx, hiprt, etype):
    """feasm_linform_dot_device fills a device vector with exactly what
    feasm_linform_dot returns in a host array."""
    if etype == O.H8:
        conn, xyz = core.gen_h8block(1.0, 1.0, 1.0, 6, 5, 4)
        xyz = core.gen_perturb(xyz, [0, 0, 0], [1, 1, 1], 0.02, 4)
        rule, order = L.RULE_GAUSS, 2
    else:
        conn, xyz = core.gen_t10block(1.0, 1.0, 1.0, 3, 3, 2, 0.02, 0.006, 4)
        rule, order = L.RULE_TET, 4
    mesh = core.DeviceMesh(ctx, etype, conn, xyz)
    mesh.set_dofs(3)
    nall = 3 * xyz.shape[0]
    force = np.array([0.5, -9.81, 2.0])
    F_host = mesh.linform_dot(rule, order, force, nall)
    buf = DeviceBuffers(hiprt)
    try:
        dF = buf.upload(np.zeros(nall))
        rc = L.lib().feasm_linform_dot_device(mesh.h, rule, order, force.ctypes.data_as(ctypes.c_void_p), 3, dF)
        assert rc == 0, L.lib().feasm_last_error(ctx.h)
        ctx.synchronize()  # the kernels run on the context's (non-blocking) stream
        F_dev = buf.download(dF, np.empty(nall))
    finally:
        buf.free()
    mesh.close()
    assert np.array_equal(F_dev, F_host)


def test_coo_assemble_many_matches_one_at_a_time(ctx):
    """feasm_coo_assemble_many buffers exactly the triples of the same
    element matrices passed one by one to feasm_coo_assemble (assemble!,
    src/AssemblyModule.jl:250-282), so makematrix! gives the same matrix, and
    it checks the DOF numbers the same way."""
    rng = np.random.default_rng(9)
    conn, _ = core.gen_h8block(1.0, 1.0, 1.0, 3, 2, 2)
    nn = int(conn.max())
    ne = conn.shape[0]
    dofs = np.stack([np.ravel(np.stack([3 * (c - 1) + k + 1 for k in range(3)], axis=1)) for c in conn])  # ne x 24
    mats = rng.normal(size=(ne, 24, 24))
    lib = L.lib()

    def run(many):
        h = ctypes.c_void_p()
        assert lib.feasm_coo_create(ctx.h, 0, ctypes.byref(h)) == 0
        assert lib.feasm_coo_start(h, 24, 24, ne, 3 * nn, 3 * nn) == 0
        if many:
            mm = np.ascontiguousarray(np.transpose(mats, (0, 2, 1)))  # each matrix column-major
            d = np.ascontiguousarray(dofs, dtype=np.int64)
            rc = lib.feasm_coo_assemble_many(h, ne, mm.ctypes.data_as(ctypes.c_void_p), d.ctypes.data_as(ctypes.c_void_p),
                                             24, d.ctypes.data_as(ctypes.c_void_p), 24)
            assert rc == 0, lib.feasm_last_error(ctx.h)
        else:
            for e in range(ne):
                mcol = np.asfortranarray(mats[e])
                d = np.ascontiguousarray(dofs[e], dtype=np.int64)
                rc = lib.feasm_coo_assemble(h, mcol.ctypes.data_as(ctypes.c_void_p), 24, 24,
                                            d.ctypes.data_as(ctypes.c_void_p), 24, d.ctypes.data_as(ctypes.c_void_p), 24)
                assert rc == 0
        assert lib.feasm_coo_ntriples(h) == ne * 576
        out = ctypes.c_void_p()
        assert lib.feasm_coo_makematrix(h, ctypes.byref(out)) == 0
        M = core.DeviceMatrix(ctx, out)
        arrs = M.csc_arrays(1)
        M.close()
        lib.feasm_coo_destroy(h)
        return arrs

    one, many = run(False), run(True)
    for a, b in zip(one, many):
        assert np.array_equal(a, b)
    # the reference's serialization: column-major per element, elements in order
    I = np.concatenate([np.repeat(dofs[e][None, :], 24, axis=0).T.ravel(order="F") for e in range(ne)])
    J = np.concatenate([np.repeat(dofs[e][:, None], 24, axis=1).T.ravel(order="F") for e in range(ne)])
    V = np.concatenate([mats[e].ravel(order="F") for e in range(ne)])
    ocp, orv, onz = O.sparse(I, J, V, 3 * nn, 3 * nn)
    assert np.array_equal(many[0], ocp) and np.array_equal(many[1], orv) and np.array_equal(many[2], onz)
    # a DOF number out of range is reported as assemble! reports it
    h = ctypes.c_void_p()
    assert lib.feasm_coo_create(ctx.h, 0, ctypes.byref(h)) == 0
    assert lib.feasm_coo_start(h, 24, 24, ne, 3 * nn, 3 * nn) == 0
    d = np.ascontiguousarray(dofs, dtype=np.int64)
    d[1, 3] = 3 * nn + 1
    mm = np.ascontiguousarray(np.transpose(mats, (0, 2, 1)))
    rc = lib.feasm_coo_assemble_many(h, ne, mm.ctypes.data_as(ctypes.c_void_p), d.ctypes.data_as(ctypes.c_void_p), 24,
                                     d.ctypes.data_as(ctypes.c_void_p), 24)
    assert rc == L.EDOF
    lib.feasm_coo_destroy(h)
