"""ctypes view of the CPU oracle (oracle/feasm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and
by the CPU-baseline leg of ``bench.py``, always as the checker.  The product
library (``finetools.jl_amd``) never imports this module.

Conventions follow the reference (Julia): node numbers, DOF numbers, row and
column indices are 1-based int64; element matrices are column-major.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

H8, T10 = 0, 1
GAUSS, TET = 0, 1
DIFFUSION_ISO, DIFFUSION_MAT, LIN_ELASTIC, DOT = 0, 1, 2, 3

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        i64 = ctypes.c_int64
        L.orc_gauss_rule3.argtypes = [ctypes.c_int, P, P]
        L.orc_tet_rule.argtypes = [ctypes.c_int, P, P]
        for n in ("orc_h8_bfun", "orc_h8_bfundpar", "orc_t10_bfun", "orc_t10_bfundpar"):
            getattr(L, n).argtypes = [P, P]
        L.orc_integration_data.argtypes = [ctypes.c_int] * 3 + [P, P, P]
        L.orc_locjac.argtypes = [ctypes.c_int, P, P, P, P, P]
        L.orc_jacobian_det.argtypes = [P]
        L.orc_jacobian_det.restype = ctypes.c_double
        L.orc_gradN.argtypes = [ctypes.c_int, P, P, P]
        L.orc_blmat3d.argtypes = [ctypes.c_int, P, P, P]
        L.orc_elmat.argtypes = [ctypes.c_int] * 5 + [P, P, P]
        L.orc_bilform_coo.argtypes = [ctypes.c_int] * 4 + [i64, P, P, ctypes.c_int, P, i64, P, P, P, P]
        L.orc_bilform_coo.restype = i64
        L.orc_bilform_coo_threads.argtypes = ([ctypes.c_int] * 4 + [i64, P, P, ctypes.c_int, P, i64, P, P, P, P]
                                              + [ctypes.c_int])
        L.orc_bilform_coo_threads.restype = i64
        L.orc_sparse.argtypes = [i64, P, P, P, i64, i64]
        L.orc_linform_dot.argtypes = [ctypes.c_int] * 3 + [i64, P, P, ctypes.c_int, P, i64, P, P]
        L.orc_hrz_lumped_coo.argtypes = [ctypes.c_int] * 3 + [i64, P, P, ctypes.c_int, P, i64, P, P, P, P]
        L.orc_hrz_lumped_coo.restype = i64
        L.orc_bilform_coo_symm.argtypes = [ctypes.c_int] * 4 + [i64, P, P, ctypes.c_int, P, i64, P, P, P, P]
        L.orc_bilform_coo_symm.restype = i64
        L.orc_symm_from_lower.argtypes = [P]
        L.orc_symm_from_lower.restype = P
        L.orc_sparse.restype = P
        L.orc_csc_nnz.argtypes = [P]
        L.orc_csc_nnz.restype = i64
        L.orc_csc_copy.argtypes = [P, P, P, P]
        L.orc_csc_free.argtypes = [P]
        L.orc_h8blockx.argtypes = [P, i64, P, i64, P, i64, P, P]
        L.orc_t4blockx_a.argtypes = [P, i64, P, i64, P, i64, P, P]
        L.orc_t4tot10.argtypes = [i64, P, i64, P, P, P]
        L.orc_t4tot10.restype = i64
        L.orc_perturb.argtypes = [i64, P, P, P, ctypes.c_double, ctypes.c_uint64]
        L.orc_min_detj.argtypes = [ctypes.c_int] * 3 + [i64, P, P]
        L.orc_min_detj.restype = ctypes.c_double
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


# ---------------------------------------------------------------- rules
def gauss_rule3(order):
    pc = np.zeros((order ** 3, 3))
    w = np.zeros(order ** 3)
    n = lib().orc_gauss_rule3(order, _p(pc), _p(w))
    if n < 0:
        raise ValueError("Gauss rule of order %d not available" % order)
    return pc, w


def tet_rule(npts):
    pc = np.zeros((max(npts, 1), 3))
    w = np.zeros(max(npts, 1))
    n = lib().orc_tet_rule(npts, _p(pc), _p(w))
    if n < 0:
        raise ValueError("Unknown number of integration points")
    return pc, w


def bfun(etype, xi):
    xi = _f64(xi)
    npe = 8 if etype == H8 else 10
    N = np.zeros(npe)
    (lib().orc_h8_bfun if etype == H8 else lib().orc_t10_bfun)(_p(xi), _p(N))
    return N


def bfundpar(etype, xi):
    xi = _f64(xi)
    npe = 8 if etype == H8 else 10
    dN = np.zeros((npe, 3))
    (lib().orc_h8_bfundpar if etype == H8 else lib().orc_t10_bfundpar)(_p(xi), _p(dN))
    return dN


def locjac(X, N, dN):
    X, N, dN = _f64(X), _f64(N), _f64(dN)
    loc = np.zeros(3)
    J = np.zeros((3, 3))
    lib().orc_locjac(X.shape[0], _p(X), _p(N), _p(dN), _p(loc), _p(J))
    return loc, J


def jacobian_det(J):
    return lib().orc_jacobian_det(_p(_f64(J)))


def gradN(dN, J):
    dN, J = _f64(dN), _f64(J)
    g = np.zeros_like(dN)
    lib().orc_gradN(dN.shape[0], _p(dN), _p(J), _p(g))
    return g


def blmat3d(gradN_, Rm=None):
    g = _f64(gradN_)
    Rm = np.eye(3) if Rm is None else _f64(Rm)
    B = np.zeros((6, 3 * g.shape[0]))
    lib().orc_blmat3d(g.shape[0], _p(g), _p(Rm), _p(B))
    return B


def rule_for(etype):
    return (GAUSS, 2) if etype == H8 else (TET, 4)


def elmat(etype, form, ndn, X, coeff, rule=None):
    """One element matrix (elmdim x elmdim, returned as a numpy array)."""
    rule, order = rule if rule is not None else rule_for(etype)
    X = _f64(X)
    c = _f64(np.atleast_1d(coeff))
    npe = X.shape[0]
    kd = ndn * npe
    Ke = np.zeros(kd * kd)
    rc = lib().orc_elmat(etype, rule, order, form, ndn, _p(X), _p(c), _p(Ke))
    if rc != 0:
        raise ValueError("bad rule")
    return Ke.reshape(kd, kd, order="F")


# ---------------------------------------------------------------- assembly
def bilform_coo(etype, form, conn, xyz, ndn, dofnums, nalldofs, coeff, rule=None, nthreads=1):
    """COO triples (I, J, V; 1-based) exactly as assemble! serializes them
    (nthreads > 1: the same triples from an OpenMP element loop)."""
    rule, order = rule if rule is not None else rule_for(etype)
    conn, xyz, dofnums = _i64(conn), _f64(xyz), _i64(dofnums)
    c = _f64(np.atleast_1d(coeff))
    ne, npe = conn.shape
    nt = ne * (ndn * npe) ** 2
    I = np.zeros(nt, np.int64)
    J = np.zeros(nt, np.int64)
    V = np.zeros(nt)
    if nthreads > 1:
        p = lib().orc_bilform_coo_threads(etype, rule, order, form, ne, _p(conn), _p(xyz), ndn, _p(dofnums),
                                          nalldofs, _p(c), _p(I), _p(J), _p(V), int(nthreads))
    else:
        p = lib().orc_bilform_coo(etype, rule, order, form, ne, _p(conn), _p(xyz), ndn, _p(dofnums),
                                  nalldofs, _p(c), _p(I), _p(J), _p(V))
    if p == -2:
        raise ValueError("Row degree of freedom < 1")
    if p == -3:
        raise ValueError("Row degree of freedom > size")
    if p < 0:
        raise ValueError("bad rule")
    return I[:p], J[:p], V[:p]


def linform_dot(etype, conn, xyz, ndn, force, dofnums=None, nalldofs=None, rule=None):
    """linform_dot(m = 3) with a constant force vector + SysvecAssembler's
    makevector!: the dense global load vector (0-based storage of 1-based DOFs)."""
    rule, order = rule if rule is not None else rule_for(etype)
    conn, xyz = _i64(conn), _f64(xyz)
    nn = xyz.shape[0]
    if dofnums is None:
        dofnums = np.arange(1, nn * ndn + 1, dtype=np.int64).reshape(nn, ndn)
        nalldofs = nn * ndn
    dofnums = _i64(dofnums)
    f = _f64(np.atleast_1d(force))
    F = np.zeros(nalldofs)
    rc = lib().orc_linform_dot(etype, rule, order, conn.shape[0], _p(conn), _p(xyz), ndn, _p(dofnums), nalldofs,
                               _p(f), _p(F))
    if rc == -2:
        raise ValueError("Row degree of freedom < 1")
    if rc == -3:
        raise ValueError("Row degree of freedom > size")
    if rc < 0:
        raise ValueError("bad rule")
    return F


def numberdofs_ebc(fixed, entperm=None):
    """numberdofs! (src/FieldModule.jl:343-377): free components numbered
    first, entities in the order entperm (1-based, default 1:n), components
    fastest; then the fixed ones.  fixed: nnodes x ndn booleans.  Returns
    (dofnums 1-based nnodes x ndn, nfreedofs)."""
    fixed = np.asarray(fixed, dtype=bool)
    n, dim = fixed.shape
    order = np.arange(n) if entperm is None else np.asarray(entperm, dtype=np.int64) - 1
    dof = np.zeros((n, dim), np.int64)
    nxt = 1
    nfree = 0
    for want in (False, True):
        for i in order:
            for j in range(dim):
                if fixed[i, j] == want:
                    dof[i, j] = nxt
                    nxt += 1
        if not want:
            nfree = nxt - 1
    return dof, nfree


def matrix_blocked_ff(colptr, rowval, nzval, nfree):
    """matrix_blocked_ff(A, nfree) (src/MatrixUtilityModule.jl:675-690): the
    leading nfree x nfree block A[1:nfree, 1:nfree] of a 1-based CSC."""
    cp = [1]
    rv, nz = [], []
    for c in range(nfree):
        for k in range(colptr[c] - 1, colptr[c + 1] - 1):
            if rowval[k] <= nfree:
                rv.append(rowval[k])
                nz.append(nzval[k])
        cp.append(len(rv) + 1)
    return np.array(cp, np.int64), np.array(rv, np.int64), np.array(nz)


def _csc_out(h, n):
    try:
        nnz = lib().orc_csc_nnz(h)
        colptr = np.zeros(n + 1, np.int64)
        rowval = np.zeros(nnz, np.int64)
        nzval = np.zeros(nnz)
        lib().orc_csc_copy(h, _p(colptr), _p(rowval), _p(nzval))
    finally:
        lib().orc_csc_free(h)
    return colptr, rowval, nzval


def symm_makematrix(I, J, V, n):
    """makematrix!(::SysmatAssemblerSparseSymm) of lower-triangle triples:
    sparse(), then S + transpose(S) keeping nonzero sums, diagonal halved."""
    I, J, V = _i64(I), _i64(J), _f64(V)
    h = lib().orc_sparse(len(I), _p(I), _p(J), _p(V), n, n)
    if not h:
        raise ValueError("index out of range")
    try:
        h2 = lib().orc_symm_from_lower(h)
    finally:
        lib().orc_csc_free(h)
    return _csc_out(h2, n)


def assemble_csc_symm(etype, form, conn, xyz, ndn, coeff, dofnums=None, nalldofs=None, rule=None):
    """bilform_* into a SysmatAssemblerSparseSymm + makematrix!."""
    rule, order = rule if rule is not None else rule_for(etype)
    conn, xyz = _i64(conn), _f64(xyz)
    nn = xyz.shape[0]
    if dofnums is None:
        dofnums = np.arange(1, nn * ndn + 1, dtype=np.int64).reshape(nn, ndn)
        nalldofs = nn * ndn
    dofnums = _i64(dofnums)
    c = _f64(np.atleast_1d(coeff))
    ne, npe = conn.shape
    kd = ndn * npe
    nt = ne * (kd * (kd + 1) // 2)
    I = np.zeros(nt, np.int64)
    J = np.zeros(nt, np.int64)
    V = np.zeros(nt)
    p = lib().orc_bilform_coo_symm(etype, rule, order, form, ne, _p(conn), _p(xyz), ndn, _p(dofnums), nalldofs,
                                   _p(c), _p(I), _p(J), _p(V))
    if p < 0:
        raise ValueError("bad rule or DOF")
    return symm_makematrix(I[:p], J[:p], V[:p], nalldofs)


def hrz_lumped_csc(etype, conn, xyz, ndn, coeff, dofnums=None, nalldofs=None, rule=None):
    """bilform_dot + SysmatAssemblerSparseHRZLumpingSymm's makematrix!: the
    lumped (diagonal) mass as CSC (colptr, rowval, nzval), 1-based."""
    rule, order = rule if rule is not None else rule_for(etype)
    conn, xyz = _i64(conn), _f64(xyz)
    nn = xyz.shape[0]
    if dofnums is None:
        dofnums = np.arange(1, nn * ndn + 1, dtype=np.int64).reshape(nn, ndn)
        nalldofs = nn * ndn
    dofnums = _i64(dofnums)
    c = _f64(np.atleast_1d(coeff))
    nt = conn.shape[0] * conn.shape[1] * ndn
    I = np.zeros(nt, np.int64)
    J = np.zeros(nt, np.int64)
    V = np.zeros(nt)
    p = lib().orc_hrz_lumped_coo(etype, rule, order, conn.shape[0], _p(conn), _p(xyz), ndn, _p(dofnums), nalldofs,
                                 _p(c), _p(I), _p(J), _p(V))
    if p == -2:
        raise ValueError("Column degree of freedom < 1")
    if p == -3:
        raise ValueError("Column degree of freedom > size")
    if p < 0:
        raise ValueError("bad rule")
    return sparse(I[:p], J[:p], V[:p], nalldofs, nalldofs)


def sparse(I, J, V, m, n):
    """CSC (colptr, rowval, nzval), 1-based, as SparseArrays.sparse(I,J,V,m,n)."""
    I, J, V = _i64(I), _i64(J), _f64(V)
    h = lib().orc_sparse(len(I), _p(I), _p(J), _p(V), m, n)
    if not h:
        raise ValueError("index out of range")
    try:
        nnz = lib().orc_csc_nnz(h)
        colptr = np.zeros(n + 1, np.int64)
        rowval = np.zeros(nnz, np.int64)
        nzval = np.zeros(nnz)
        lib().orc_csc_copy(h, _p(colptr), _p(rowval), _p(nzval))
    finally:
        lib().orc_csc_free(h)
    return colptr, rowval, nzval


def assemble_csc(etype, form, conn, xyz, ndn, coeff, dofnums=None, nalldofs=None, rule=None):
    """bilform_* + makematrix! with node-major numbering (numberdofs!, no EBCs)."""
    nn = xyz.shape[0]
    if dofnums is None:
        dofnums = np.arange(1, nn * ndn + 1, dtype=np.int64).reshape(nn, ndn)
        nalldofs = nn * ndn
    I, J, V = bilform_coo(etype, form, conn, xyz, ndn, dofnums, nalldofs, coeff, rule)
    return sparse(I, J, V, nalldofs, nalldofs)


# ---------------------------------------------------------------- meshes
def linearspace(start, stop, n):
    """Grid coordinates (stop - start) * i / (n - 1) + start, the formula the
    product's generators use as well (exact for [0, 1] grids)."""
    i = np.arange(n, dtype=np.float64)
    return start + (stop - start) * i / (n - 1)


def h8blockx(xs, ys, zs):
    xs, ys, zs = _f64(xs), _f64(ys), _f64(zs)
    ne = (len(xs) - 1) * (len(ys) - 1) * (len(zs) - 1)
    nn = len(xs) * len(ys) * len(zs)
    conn = np.zeros((ne, 8), np.int64)
    xyz = np.zeros((nn, 3))
    lib().orc_h8blockx(_p(xs), len(xs), _p(ys), len(ys), _p(zs), len(zs), _p(conn), _p(xyz))
    return conn, xyz


def h8block(L, W, H, nL, nW, nH):
    return h8blockx(linearspace(0.0, L, nL + 1), linearspace(0.0, W, nW + 1), linearspace(0.0, H, nH + 1))


def t4blockx_a(xs, ys, zs):
    xs, ys, zs = _f64(xs), _f64(ys), _f64(zs)
    ne = 6 * (len(xs) - 1) * (len(ys) - 1) * (len(zs) - 1)
    nn = len(xs) * len(ys) * len(zs)
    conn = np.zeros((ne, 4), np.int64)
    xyz = np.zeros((nn, 3))
    lib().orc_t4blockx_a(_p(xs), len(xs), _p(ys), len(ys), _p(zs), len(zs), _p(conn), _p(xyz))
    return conn, xyz


def t4tot10(conn4, xyz4):
    conn4, xyz4 = _i64(conn4), _f64(xyz4)
    ne = conn4.shape[0]
    conn10 = np.zeros((ne, 10), np.int64)
    nn = lib().orc_t4tot10(xyz4.shape[0], _p(xyz4), ne, _p(conn4), _p(conn10), None)
    xyz10 = np.zeros((nn, 3))
    lib().orc_t4tot10(xyz4.shape[0], _p(xyz4), ne, _p(conn4), _p(conn10), _p(xyz10))
    return conn10, xyz10


def t10blockx_a(xs, ys, zs):
    c4, x4 = t4blockx_a(xs, ys, zs)
    return t4tot10(c4, x4)


def perturb(xyz, lo, hi, amp, seed):
    xyz = _f64(xyz).copy()
    lo, hi = _f64(lo), _f64(hi)
    lib().orc_perturb(xyz.shape[0], _p(xyz), _p(lo), _p(hi), amp, seed)
    return xyz


def min_detj(etype, conn, xyz, rule=None):
    rule, order = rule if rule is not None else rule_for(etype)
    conn, xyz = _i64(conn), _f64(xyz)
    return lib().orc_min_detj(etype, rule, order, conn.shape[0], _p(conn), _p(xyz))


def block_mesh(etype, n, L=1.0, amp_frac=0.0, seed=0):
    """Perturbed unit block: H8 n^3 or T10 (6 tets/cell) n^3, interior nodes
    moved by U[-1,1) * amp_frac * h."""
    xs = linearspace(0.0, L, n[0] + 1)
    ys = linearspace(0.0, L, n[1] + 1)
    zs = linearspace(0.0, L, n[2] + 1)
    if etype == H8:
        conn, xyz = h8blockx(xs, ys, zs)
    else:
        conn, xyz = t10blockx_a(xs, ys, zs)
    if amp_frac:
        h = L / max(n)
        xyz = perturb(xyz, [0.0, 0.0, 0.0], [L, L, L], amp_frac * h, seed)
    return conn, xyz


def iso_D(E, nu):
    """6x6 isotropic elasticity matrix in the strain order xx, yy, zz, xy, xz, yz
    (engineering shear), as the configurations specify it."""
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    for i in range(3):
        D[i, i] = lam + 2 * mu
        D[3 + i, 3 + i] = mu
    return D
