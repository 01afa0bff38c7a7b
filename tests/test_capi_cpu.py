"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, its host-side mesh generators reproduce the oracle's, and the
assembler protocol raises the reference's errors."""
import ctypes
import os
import re

import numpy as np
import pytest

import feasm
from feasm import _lib as L
from feasm import core
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "feasm.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(feasm_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    syms = declared_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(lib, s), s
    # and the Python binding types every one of them
    assert set(syms) == set(L.SIGNATURES)


def test_h8_generator_matches_oracle():
    for n in ((1, 1, 1), (10, 10, 10), (3, 5, 2)):
        c, x = core.gen_h8block(1.0, 1.0, 1.0, *n)
        oc, ox = O.block_mesh(O.H8, n)
        assert np.array_equal(c, oc) and np.array_equal(x, ox)
    c, x = core.gen_h8block(2.0, 3.0, 0.5, 2, 3, 4)
    oc, ox = O.h8block(2.0, 3.0, 0.5, 2, 3, 4)
    assert np.array_equal(c, oc) and np.array_equal(x, ox)


def test_perturb_matches_oracle():
    c, x = core.gen_h8block(1.0, 1.0, 1.0, 7, 6, 5)
    a = core.gen_perturb(x, [0, 0, 0], [1, 1, 1], 0.15 / 7, 99)
    b = O.perturb(x, [0, 0, 0], [1, 1, 1], 0.15 / 7, 99)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("n,av,am", [((5, 10, 5), 0.0, 0.0), ((4, 3, 5), 0.1 / 5, 0.03 / 5)])
def test_t10_generator_matches_oracle(n, av, am):
    c, x = core.gen_t10block(1.0, 1.0, 1.0, *n, av, am, 17)
    xs, ys, zs = (O.linearspace(0.0, 1.0, k + 1) for k in n)
    c4, x4 = O.t4blockx_a(xs, ys, zs)
    if av:
        x4 = O.perturb(x4, [0, 0, 0], [1, 1, 1], av, 17)
    c10, x10 = O.t4tot10(c4, x4)
    if am:
        nv = x4.shape[0]
        x10[nv:] = O.perturb(x10[nv:], [0, 0, 0], [1, 1, 1], am, 18)
    assert np.array_equal(c, c10) and np.array_equal(x, x10)
    if n == (5, 10, 5):
        assert x.shape[0] == 2541 and c.shape[0] == 1500


def _host_assembler():
    h = ctypes.c_void_p()
    assert L.lib().feasm_coo_create(None, 0, ctypes.byref(h)) == 0
    return h


def test_assembler_protocol_errors_like_reference():
    lib = L.lib()
    a = _host_assembler()
    try:
        m = np.asfortranarray(np.ones((3, 4)))
        r = np.array([1, 7, 5], np.int64)
        c = np.array([5, 2, 1, 4], np.int64)
        # assemble! before startassembly!
        assert lib.feasm_coo_assemble(a, L.ptr(m), 3, 4, L.ptr(r), 3, L.ptr(c), 4) == L.ESTATE
        assert lib.feasm_coo_start(a, 5, 5, 3, 7, 7) == 0
        assert lib.feasm_coo_assemble(a, L.ptr(m), 3, 4, L.ptr(r), 3, L.ptr(c), 4) == 0
        assert lib.feasm_coo_ntriples(a) == 12
        # "Wrong size of matrix" (src/AssemblyModule.jl:265)
        assert lib.feasm_coo_assemble(a, L.ptr(m), 3, 4, L.ptr(r), 2, L.ptr(c), 4) == L.ESIZE
        assert lib.feasm_last_error(None) == b"Wrong size of matrix"
        for bad, msg in ((np.array([0, 2, 1, 4]), b"Column degree of freedom < 1"),
                         (np.array([5, 8, 1, 4]), b"Column degree of freedom > size")):
            bad = bad.astype(np.int64)
            assert lib.feasm_coo_assemble(a, L.ptr(m), 3, 4, L.ptr(r), 3, L.ptr(bad), 4) == L.EDOF
            assert lib.feasm_last_error(None) == msg
        for bad, msg in ((np.array([1, 0, 5]), b"Row degree of freedom < 1"),
                         (np.array([1, 9, 5]), b"Row degree of freedom > size")):
            bad = bad.astype(np.int64)
            assert lib.feasm_coo_assemble(a, L.ptr(m), 3, 4, L.ptr(bad), 3, L.ptr(c), 4) == L.EDOF
            assert lib.feasm_last_error(None) == msg
        # failed calls add nothing; buffers grow by nrows*ncols*1000 when full
        assert lib.feasm_coo_ntriples(a) == 12
        for _ in range(10):
            assert lib.feasm_coo_assemble(a, L.ptr(m), 3, 4, L.ptr(r), 3, L.ptr(c), 4) == 0
        assert lib.feasm_coo_ntriples(a) == 132
        # no device context: makematrix! reports it
        out = ctypes.c_void_p()
        assert lib.feasm_coo_makematrix(a, ctypes.byref(out)) == L.ESTATE
        # nomatrixresult keeps the triples and returns no matrix
        assert lib.feasm_coo_set_nomatrixresult(a, 1) == 0
        assert lib.feasm_coo_makematrix(a, ctypes.byref(out)) == 0 and not out.value
        assert lib.feasm_coo_ntriples(a) == 132
    finally:
        lib.feasm_coo_destroy(a)


def test_context_without_gpu_fails_loudly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(feasm.FeasmError):
        feasm.Context(0)


def test_numberdofs_node_major():
    from feasm.finetools import NodalField, nalldofs, numberdofs

    u = NodalField(np.zeros((4, 3)))
    numberdofs(u)
    assert u.dofnums.tolist() == [[1, 2, 3], [4, 5, 6], [7, 8, 9], [10, 11, 12]]
    assert nalldofs(u) == 12
    numberdofs(u, [4, 3, 2, 1])
    assert u.dofnums[3].tolist() == [1, 2, 3]


def _host_assembler_kind(kind):
    h = ctypes.c_void_p()
    assert L.lib().feasm_coo_create_kind(None, kind, 0, ctypes.byref(h)) == 0
    return h


def test_symm_diag_hrz_protocols_buffer_and_fail_like_reference():
    """Host side of SysmatAssemblerSparseSymm / Diag / HRZLumpingSymm
    (src/AssemblyModule.jl:424-549, :671-760, :1031-1114): buffer sizes, the
    number of triples each assemble! adds, growth, and the error texts."""
    lib = L.lib()
    m = np.asfortranarray(np.arange(16, dtype=np.float64).reshape(4, 4))
    d = np.array([5, 2, 1, 4], np.int64)
    for kind, per in ((L.ASM_SYMM, 10), (L.ASM_DIAG, 4), (L.ASM_HRZ, 4)):
        a = _host_assembler_kind(kind)
        try:
            if kind != L.ASM_DIAG:  # square element matrices only
                assert lib.feasm_coo_start(a, 4, 3, 2, 7, 7) == L.ESIZE
                want = (b"Symmetric sparse matrix is assumed to be assembled from square matrices"
                        if kind == L.ASM_SYMM else
                        b"Diagonal sparse matrix is assumed to be assembled from square matrices")
                assert lib.feasm_last_error(None) == want
            assert lib.feasm_coo_start(a, 4, 4, 2, 7, 6) == L.EINVAL
            assert lib.feasm_last_error(None) == b"Row and column info do not agree"
            b = _host_assembler_kind(kind)
            assert lib.feasm_coo_start(b, 4, 4, 2, 7, 7) == 0
            assert lib.feasm_coo_assemble(b, L.ptr(m), 4, 4, L.ptr(d), 4, L.ptr(d), 4) == 0
            assert lib.feasm_coo_ntriples(b) == per
            # rows and columns of different lengths
            assert lib.feasm_coo_assemble(b, L.ptr(m), 4, 4, L.ptr(d), 3, L.ptr(d), 4) == L.ESIZE
            assert lib.feasm_last_error(None) == b"Size mismatch"
            bad = np.array([5, 2, 9, 4], np.int64)
            assert lib.feasm_coo_assemble(b, L.ptr(m), 4, 4, L.ptr(bad), 4, L.ptr(bad), 4) == L.EDOF
            # Symm meets the rows of the first column before the third column
            assert lib.feasm_last_error(None) == (b"Row degree of freedom > size" if kind == L.ASM_SYMM
                                                  else b"Column degree of freedom > size")
            bad = np.array([0, 2, 1, 4], np.int64)
            assert lib.feasm_coo_assemble(b, L.ptr(m), 4, 4, L.ptr(bad), 4, L.ptr(bad), 4) == L.EDOF
            assert lib.feasm_last_error(None) == b"Column degree of freedom < 1"
            for _ in range(5):  # grows past the 2-element buffer
                assert lib.feasm_coo_assemble(b, L.ptr(m), 4, 4, L.ptr(d), 4, L.ptr(d), 4) == 0
            assert lib.feasm_coo_ntriples(b) == 6 * per
            lib.feasm_coo_destroy(b)
        finally:
            lib.feasm_coo_destroy(a)
    # the symmetric assembler checks every row of the lower triangle, the
    # diagonal ones only their columns
    a = _host_assembler_kind(L.ASM_SYMM)
    b = _host_assembler_kind(L.ASM_DIAG)
    try:
        for h in (a, b):
            assert lib.feasm_coo_start(h, 4, 4, 1, 7, 7) == 0
        rows = np.array([5, 2, 1, 8], np.int64)
        cols = np.array([5, 2, 1, 4], np.int64)
        assert lib.feasm_coo_assemble(a, L.ptr(m), 4, 4, L.ptr(rows), 4, L.ptr(cols), 4) == L.EDOF
        assert lib.feasm_last_error(None) == b"Row degree of freedom > size"
        assert lib.feasm_coo_assemble(b, L.ptr(m), 4, 4, L.ptr(rows), 4, L.ptr(cols), 4) == 0
    finally:
        lib.feasm_coo_destroy(a)
        lib.feasm_coo_destroy(b)
