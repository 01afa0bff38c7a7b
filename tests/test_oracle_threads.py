"""The oracle's OpenMP element loop (used only for the multi-core CPU
baseline) writes exactly the serial loop's triples, in the same order."""
import numpy as np

import oracle as O


def test_threaded_element_loop_matches_serial():
    for et, n in ((O.H8, (6, 5, 4)), (O.T10, (2, 3, 2))):
        conn, xyz = O.block_mesh(et, n, L=1.0, amp_frac=0.1, seed=4)
        nn = xyz.shape[0]
        dofnums = np.arange(1, 3 * nn + 1, dtype=np.int64).reshape(nn, 3)
        D = O.iso_D(70e3, 0.33).ravel()
        a = O.bilform_coo(et, O.LIN_ELASTIC, conn, xyz, 3, dofnums, 3 * nn, D)
        b = O.bilform_coo(et, O.LIN_ELASTIC, conn, xyz, 3, dofnums, 3 * nn, D, nthreads=4)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
