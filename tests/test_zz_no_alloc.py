"""GPU test of feasm_sparse_device_noalloc, the entry point that allocates no
device memory: its CSC equals feasm_sparse_device's bitwise on two calls in a
row (the second smaller, so stale scratch contents would show), and both
equal the oracle's sparse(); the size limit and the index range check are
refused from the host or by a flag, before any output is written."""
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


class Buffers:
    def __init__(self, hip):
        self.hip, self.bufs = hip, []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(int(nbytes), 8)) == 0
        self.bufs.append(p)
        return p

    def upload(self, a):
        p = self.alloc(a.nbytes)
        if a.nbytes:
            assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        return p

    def download(self, p, out):
        if out.nbytes:
            assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def triples(rng, m, n, nt):
    I = rng.integers(0, m, nt).astype(np.int32)
    J = rng.integers(0, n, nt).astype(np.int32)
    V = rng.normal(size=nt)
    V[::13] = 0.0                # explicit zeros are kept
    if nt > 10:                  # an order-sensitive duplicate: folds to 0 left to right
        I[7:10], J[7:10], V[7:10] = I[7], J[7], [1e16, 1.0, -1e16]
    return I, J, V


def noalloc(ctx, buf, I, J, V, m, n):
    nt = len(V)
    dI, dJ, dV = buf.upload(I), buf.upload(J), buf.upload(V)
    dcp, drv, dnz = buf.alloc(8 * (n + 1)), buf.alloc(4 * nt), buf.alloc(8 * nt)
    nnz = ctypes.c_int64(-1)
    rc = L.lib().feasm_sparse_device_noalloc(ctx.h, nt, dI, dJ, dV, m, n, dcp, drv, dnz, ctypes.byref(nnz))
    assert rc == L.OK, L.lib().feasm_last_error(ctx.h)
    k = nnz.value
    cp = buf.download(dcp, np.empty(n + 1, np.int64))
    rv = buf.download(drv, np.empty(k, np.int32))
    nz = buf.download(dnz, np.empty(k, np.float64))
    return cp, rv.astype(np.int64), nz


def allocating(ctx, buf, I, J, V, m, n):
    dI, dJ, dV = buf.upload(I), buf.upload(J), buf.upload(V)
    h = ctypes.c_void_p()
    rc = L.lib().feasm_sparse_device(ctx.h, len(V), dI, dJ, dV, m, n, ctypes.byref(h))
    assert rc == L.OK, L.lib().feasm_last_error(ctx.h)
    M = core.DeviceMatrix(ctx, h)
    out = M.csc_arrays(0)
    M.close()
    return out


def test_noalloc_matches_the_allocating_entry_point(ctx, hiprt):
    rng = np.random.default_rng(77)
    buf = Buffers(hiprt)
    try:
        # two calls in a row, the second smaller; then the maximum size; then empty
        for m, n, nt in ((3001, 2003, 900_000), (97, 61, 5_000), (20_000, 20_000, 1_048_576),
                         (40, 30, 0)):
            say("noalloc sparse: m %d n %d triples %d" % (m, n, nt))
            I, J, V = triples(rng, m, n, nt)
            a = noalloc(ctx, buf, I, J, V, m, n)
            b = allocating(ctx, buf, I, J, V, m, n)
            for x, y in zip(a, b):
                assert np.array_equal(x, y)
            if nt <= 5_000:
                ocp, orv, onz = O.sparse(I.astype(np.int64) + 1, J.astype(np.int64) + 1, V, m, n)
                assert np.array_equal(a[0], ocp - 1) and np.array_equal(a[1], orv - 1)
                assert np.array_equal(a[2], onz)
            buf.free()
    finally:
        buf.free()


def test_noalloc_refuses_beyond_its_limits(ctx, hiprt):
    buf = Buffers(hiprt)
    try:
        nt = 1_048_577
        say("noalloc sparse: one triple over the limit")
        dI, dJ, dV = buf.alloc(4 * nt), buf.alloc(4 * nt), buf.alloc(8 * nt)
        dcp, drv, dnz = buf.alloc(8 * 11), buf.alloc(4 * nt), buf.alloc(8 * nt)
        nnz = ctypes.c_int64(-1)
        rc = L.lib().feasm_sparse_device_noalloc(ctx.h, nt, dI, dJ, dV, 10, 10, dcp, drv, dnz, ctypes.byref(nnz))
        assert rc == L.ENOMEM and b"FEASM_NOALLOC_MAX_TRIPLES" in L.lib().feasm_last_error(ctx.h)
        buf.free()
        say("noalloc sparse: an index out of range")
        rng = np.random.default_rng(5)
        I, J, V = triples(rng, 50, 40, 3000)
        I[2999] = 50
        dI, dJ, dV = buf.upload(I), buf.upload(J), buf.upload(V)
        dcp, drv, dnz = buf.alloc(8 * 41), buf.alloc(4 * 3000), buf.alloc(8 * 3000)
        rc = L.lib().feasm_sparse_device_noalloc(ctx.h, 3000, dI, dJ, dV, 50, 40, dcp, drv, dnz,
                                                 ctypes.byref(nnz))
        assert rc == L.EDOF
        buf.free()
        # the scratch still serves a good call afterwards
        I[2999] = 49
        a = noalloc(ctx, buf, I, J, V, 50, 40)
        b = allocating(ctx, buf, I, J, V, 50, 40)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    finally:
        buf.free()
