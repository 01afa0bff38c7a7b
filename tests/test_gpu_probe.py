"""GPU test of the element-Gram measurement helper: the FP64 matrix-core
variant (v_mfma_f64_16x16x4f64, with the lane layout the kernel assumes)
gives the same node-pair blocks as the FP64 vector variant -- bitwise, the
products and their order coincide -- on 64 elements, and the same checksum
to rounding over 4096 elements."""
import ctypes

import numpy as np
import pytest

from feasm import _lib as L

pytestmark = pytest.mark.gpu


def run(ctx, v, ne, reps, keep):
    ms, cs = ctypes.c_double(), ctypes.c_double()
    g = np.zeros(max(keep, 1) * 324)
    rc = L.lib().feasm_probe_gram(ctx.h, v, ne, reps, ctypes.byref(ms), ctypes.byref(cs), L.ptr(g) if keep else None,
                                  keep)
    assert rc == L.OK, L.lib().feasm_last_error(ctx.h)
    return cs.value, g[:keep * 324]


def test_matrix_core_gram_matches_vector_gram(ctx):
    c0, g0 = run(ctx, 0, 1024, 1, 64)
    c1, g1 = run(ctx, 1, 1024, 1, 64)
    assert np.count_nonzero(g0) > 0.9 * g0.size
    assert np.array_equal(g0, g1)
    c0, _ = run(ctx, 0, 4096, 2, 0)
    c1, _ = run(ctx, 1, 4096, 2, 0)
    assert abs(c0 - c1) <= 1e-12 * abs(c0)
