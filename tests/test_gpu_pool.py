"""The device memory pool: released buffers are kept for reuse by later
calls and contexts, and feasm_pool_trim frees them (the library keeps
working afterwards, with the same results)."""
import numpy as np
import pytest

from feasm import _lib as L
from feasm import core
import oracle as O

pytestmark = pytest.mark.gpu


def assemble(ctx):
    conn, xyz = core.gen_h8block(1.0, 1.0, 1.0, 5, 4, 3)
    m = core.DeviceMesh(ctx, L.H8, conn, xyz)
    m.set_dofs(3)
    M = m.bilform(L.LIN_ELASTIC, L.RULE_GAUSS, 2, np.ravel(O.iso_D(200e3, 0.3)))
    out = M.csc_arrays(base=1)
    M.close()
    m.close()
    return out


def test_pool_trim(ctx):
    first = assemble(ctx)
    freed = core.pool_trim(0)
    print("freed", freed, flush=True)
    assert freed > 0
    assert core.pool_trim(0) == 0
    # buffers kept from a context that has ended serve a new one
    c2 = core.Context(0)
    try:
        again = assemble(c2)
    finally:
        c2.close()
    for x, y in zip(first, again):
        assert np.array_equal(x, y)
    assert core.pool_trim(0) > 0
