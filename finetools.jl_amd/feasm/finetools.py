"""Host-side mirror of FinEtools.jl's assembly interface, backed by libfeasm.

Names, argument meaning and error behaviour follow the reference so that code
written against it reads the same (Julia's trailing ``!`` cannot appear in
Python names: ``startassembly!`` is ``startassembly`` and so on):

  FENodeSet, FESetH8, FESetT10      src/FENodeSetModule.jl:26-32, src/FESetModule.jl:953,1393
  NodalField, numberdofs, nalldofs  src/NodalFieldModule.jl:20-62, src/FieldModule.jl:102,328-377
  GaussRule, TetRule                src/IntegRuleModule.jl:206-398, 483-516
  IntegDomain, FEMMBase             src/IntegDomainModule.jl:43-61, src/FEMMBaseModule.jl:72-84
  DataCache, DeforModelRed3D        src/DataCacheModule.jl:129-175, src/DeforModelRedModule.jl:64
  SysmatAssemblerSparse             src/AssemblyModule.jl:88-329
    startassembly, assemble, makematrix, setnomatrixresult, expectedntriples
  bilform_diffusion                 src/FEMMBaseModule.jl:1462-1545
  bilform_lin_elastic               src/FEMMBaseModule.jl:1774-1824
  bilform_dot                       src/FEMMBaseModule.jl:1335-1376
  H8block, T10block                 src/MeshHexahedronModule.jl:39-52, src/MeshTetrahedronModule.jl:249-261

Matrices come back as scipy.sparse.csc_matrix (0-based).  A bilform_* call
with a SysmatAssemblerSparse runs the whole element loop on the GPU (the
device keeps mesh, DOFs and the CSC pattern resident between calls); other
calls of the generic protocol buffer triples on the host and build the CSC
with the device sparse().
"""
import ctypes

import numpy as np
import scipy.sparse as sp

from . import _lib as L
from ._lib import check, lib, ptr
from .core import DeviceMatrix, DeviceMesh, default_context, gen_h8block, gen_t10block


# ------------------------------------------------------------------ mesh data
class FENodeSet:
    def __init__(self, xyz):
        self.xyz = np.array(xyz, dtype=np.float64, copy=True)

    def count(self):
        return self.xyz.shape[0]


class _FESet:
    NPE = 0
    ETYPE = -1

    def __init__(self, conn):
        conn = np.asarray(conn, dtype=np.int64)
        if conn.ndim != 2 or conn.shape[1] != self.NPE:
            raise ValueError("Wrong number of nodes per element")
        self.conn = np.ascontiguousarray(conn)
        self.label = np.zeros(conn.shape[0], np.int64)

    def count(self):
        return self.conn.shape[0]

    def nodesperelem(self):
        return self.NPE


class FESetH8(_FESet):
    NPE = 8
    ETYPE = L.H8


class FESetT10(_FESet):
    NPE = 10
    ETYPE = L.T10


def H8block(Length, Width, Height, nL, nW, nH):
    conn, xyz = gen_h8block(float(Length), float(Width), float(Height), int(nL), int(nW), int(nH))
    return FENodeSet(xyz), FESetH8(conn)


def T10block(Length, Width, Height, nL, nW, nH, orientation="a"):
    if orientation not in ("a", ":a"):
        raise ValueError("only orientation :a is generated")
    conn, xyz = gen_t10block(float(Length), float(Width), float(Height), int(nL), int(nW), int(nH))
    return FENodeSet(xyz), FESetT10(conn)


class FESetT4(_FESet):
    """Linear tetrahedra: read by the importers; the device path assembles H8
    and T10 meshes only (a T4 set raises at assembly)."""
    NPE = 4
    ETYPE = -1


# ------------------------------------------------------------------ mesh import
def _fixupdecimal(s):
    """A NASTRAN number with the exponent letter left out ("-1.66618812195+1")
    made parseable: an E goes before every sign that follows a digit
    (src/MeshImportModule.jl:37-49)."""
    out = []
    for i, ch in enumerate(s):
        if ch in "+-" and i > 0 and s[i - 1].isdigit():
            out.append("E")
        out.append(ch)
    return "".join(out)


def import_NASTRAN(filename, allocationchunk=None, expectfixedformat=False):
    """GRID and CTETRA records of a NASTRAN bulk-data file
    (src/MeshImportModule.jl:69-196): free-format GRID, or large-field fixed
    format GRID* with its continuation line; CTETRA with 4 or 10 nodes,
    continued on the next line when the record is cut (fixed format: a
    72-column record).  Nodes must be numbered serially.  One finite element
    set per property id, in order of first appearance.  As in the reference
    the file's last line is not examined (it is ENDDATA).  allocationchunk is
    accepted for signature compatibility (storage grows as needed)."""
    del allocationchunk
    with open(filename) as f:
        lines = [ln.rstrip("\r\n") for ln in f]
    nodes, elems = [], []
    cur = 0
    while cur < len(lines) - 1:
        temp = lines[cur]
        cur += 1
        if len(temp) >= 4 and temp[:4].upper() == "GRID":
            fixed, large = expectfixedformat, False
            if len(temp) >= 5 and temp[:5].upper() == "GRID*":
                large, fixed = True, True
            if fixed and not large:
                raise ValueError("Can handle either free format or large-field fixed format")
            if fixed:
                nid = float(_fixupdecimal(temp[8:24]))
                x = float(_fixupdecimal(temp[40:56]))
                y = float(_fixupdecimal(temp[56:72]))
                temp = lines[cur]
                cur += 1
                z = float(_fixupdecimal(temp[8:24]))
                nodes.append((nid, x, y, z))
            else:
                a = temp.replace(",", " ").split()
                nodes.append(tuple(float(a[k]) for k in range(1, 5)))
        if len(temp) >= 6 and temp[:6].upper() == "CTETRA":
            continuation = ""
            fixed = len(temp) == 72 or expectfixedformat
            if fixed:
                continuation = temp[min(len(temp), 72) - 1:]
                temp = temp[:min(len(temp), 72)]
            a = temp.replace(",", " ").split()
            eid, pid = int(a[1]), int(a[2])
            if len(a) == 7:
                nper = 4
            else:
                nper = 10
                if len(a) < 13:  # the record continues on the next line
                    temp = lines[cur]
                    cur += 1
                    if fixed:
                        temp = temp[len(continuation):min(len(temp), 72)]
                    a = a + temp.strip().replace(",", " ").split()
            elems.append((eid, pid, [int(a[k]) for k in range(3, 3 + nper)]))
    node = np.array(nodes, dtype=np.float64).reshape(-1, 4)
    if not np.array_equal(node[:, 0], np.arange(1, node.shape[0] + 1)):
        raise ValueError("Nodes are not in serial order")
    npers = sorted({len(e[2]) for e in elems})
    if len(npers) != 1:
        raise ValueError("Cannot handle mixture of element types")
    conn = np.array([e[2] for e in elems], dtype=np.int64)
    pids = [e[1] for e in elems]
    upids = list(dict.fromkeys(pids))
    cls = FESetT4 if npers[0] == 4 else FESetT10
    fesets = [cls(conn[[i for i, q in enumerate(pids) if q == pid]]) for pid in upids]
    return {"fens": FENodeSet(node[:, 1:4]), "fesets": fesets, "property_ids": upids}


def import_ABAQUS(filename, allocationchunk=None):
    """*NODE, *ELEMENT, *NSET and *ELSET sections of an Abaqus input file
    (src/MeshImportModule.jl:236-470): nodes renumbered serially in file
    order and the connectivity with them; consecutive *ELEMENT sections are
    read (C3D8 -> H8, C3D10 -> T10, C3D4/DC3D4 -> T4; other types are
    reported in "warnings", like every type the device path does not
    assemble); an element line ending in a comma continues on the next line;
    GENERATE ranges expand.  nsets hold the new node numbers."""
    del allocationchunk
    with open(filename) as f:
        lines = [ln.strip().upper() for ln in f]
    warnings = []
    # nodes: the first *NODE section
    node = []
    i = 0
    while i < len(lines):
        if lines[i][:5] == "*NODE":
            i += 1
            while i < len(lines) and lines[i][:1] != "*":
                node.append([float(v) for v in lines[i].split(",") if v.strip() != ""][:4])
                i += 1
            break
        i += 1
    node = np.array(node, dtype=np.float64).reshape(-1, 4)
    ids = node[:, 0].astype(np.int64)
    newnum = np.zeros(int(ids.max()) + 1, dtype=np.int64)
    newnum[ids] = np.arange(1, len(ids) + 1)
    # element sections: consecutive *ELEMENT blocks from the first one
    sections = []
    i = 0
    while i < len(lines) and lines[i][:8] != "*ELEMENT":
        i += 1
    while i < len(lines) and lines[i][:8] == "*ELEMENT":
        header, rows = lines[i], []
        i += 1
        while i < len(lines) and lines[i][:1] != "*":
            a = lines[i].split(",")
            i += 1
            if a[-1] == "" and i < len(lines):
                a = a[:-1] + lines[i].split(",")
                i += 1
            rows.append([int(v) for v in a if v.strip() != ""])
        sections.append((header, rows))
    fesets = []
    for header, rows in sections:
        etype = None
        for part in header.split(","):
            kv = part.split("=")
            if kv[0].strip() == "TYPE" and len(kv) > 1:
                etype = kv[1].strip()
        if etype is not None and etype[:4] == "C3D8":
            cls, n = FESetH8, 8
        elif etype is not None and etype[:5] == "C3D10":
            cls, n = FESetT10, 10
        elif etype is not None and (etype[:4] == "C3D4" or etype[:5] == "DC3D4"):
            cls, n = FESetT4, 4
        else:
            warnings.append("Don't know how to handle " + header)
            continue
        conn = np.array([r[1:1 + n] for r in rows], dtype=np.int64).reshape(-1, n)
        fesets.append(cls(newnum[conn]))
    # node and element sets
    def sets(tag, key):
        out = {}
        j = 0
        while j < len(lines):
            if lines[j][:len(tag)] == tag:
                name, gen = "", False
                for part in lines[j].split(","):
                    kv = [t.strip() for t in part.split("=")]
                    if kv[0] == key and len(kv) > 1:
                        name = kv[1]
                    if kv[0] == "GENERATE":
                        gen = True
                vals = []
                j += 1
                while j < len(lines) and lines[j][:1] != "*":
                    a = [int(v) for v in lines[j].split(",") if v.strip() != ""]
                    vals.extend(range(a[0], a[1] + 1, a[2]) if gen else a)
                    j += 1
                out[name] = np.array(vals, dtype=np.int64)
                continue
            j += 1
        return out
    nsets = {k: newnum[v] for k, v in sets("*NSET", "NSET").items()}
    elsets = sets("*ELSET", "ELSET")
    return {"fens": FENodeSet(node[:, 1:4]), "fesets": fesets, "nsets": nsets, "elsets": elsets,
            "warnings": warnings}


# ------------------------------------------------------------------ fields
class NodalField:
    """values (nnodes x ndn), the fixed flags set by setebc (DOF_KIND_DATA in
    the reference) and dofnums filled by numberdofs."""

    def __init__(self, values):
        v = np.asarray(values, dtype=np.float64)
        if v.ndim == 1:
            v = v.reshape(-1, 1)
        self.values = v.copy()
        self.dofnums = np.zeros(v.shape, np.int64)
        self.fixed = np.zeros(v.shape, bool)
        self.nfreedofs = 0
        self.nalldofs_ = 0
        self._entperm_default = True

    def ndofs(self):
        return self.values.shape[1]

    def nents(self):
        return self.values.shape[0]


def setebc(u, fenids, is_fixed=True, comp=1, val=0.0):
    """setebc!(self, fenids, is_fixed, comp, val) (src/FieldModule.jl:400-455):
    1-based node ids and component; a fixed DOF takes the value, a released
    one is reset to zero."""
    fenids = np.atleast_1d(np.asarray(fenids, dtype=np.int64))
    val = np.broadcast_to(np.asarray(val, dtype=np.float64), fenids.shape)
    if not (1 <= comp <= u.values.shape[1]):
        raise ValueError("Requested  nonexistent  degree of freedom")
    if fenids.size and (fenids.max() > u.values.shape[0] or fenids.min() < 1):
        raise ValueError("Requested nonexistent node")
    for i, v in zip(fenids, val):
        u.fixed[i - 1, comp - 1] = bool(is_fixed)
        u.values[i - 1, comp - 1] = v if is_fixed else 0.0
    return u


def numberdofs(u, entperm=None):
    """numberdofs! (src/FieldModule.jl:343-377): the free components numbered
    consecutively (entities in the order entperm, default 1:n; components
    fastest), then the fixed ones, again consecutively."""
    n, dim = u.values.shape
    order = np.arange(n) if entperm is None else np.asarray(entperm, dtype=np.int64) - 1
    u._entperm_default = entperm is None or np.array_equal(order, np.arange(n))
    nxt = 1
    for want_fixed in (False, True):
        for i in order:
            for j in range(dim):
                if u.fixed[i, j] == want_fixed:
                    u.dofnums[i, j] = nxt
                    nxt += 1
        if not want_fixed:
            u.nfreedofs = nxt - 1
    u.nalldofs_ = nxt - 1
    return u


def nalldofs(u):
    return int(u.nalldofs_)


def nfreedofs(u):
    return int(u.nfreedofs)


# ------------------------------------------------------------------ quadrature + FEMM
class GaussRule:
    def __init__(self, dim=1, order=1):
        if not (1 <= dim <= 3):
            raise ValueError("Gauss rule of dimension %d not available" % dim)
        if order not in (1, 2, 3, 4):
            raise ValueError("Gauss rule of order %d not available" % order)
        self.dim, self.order, self.npts = dim, order, order ** dim
        self.kind = L.RULE_GAUSS


class TetRule:
    def __init__(self, npts=1):
        if npts not in (1, 4, 5):
            raise ValueError("Unknown number of integration points")
        self.npts = npts
        self.order = npts
        self.kind = L.RULE_TET


class IntegDomain:
    def __init__(self, fes, integration_rule, otherdimension=1.0, axisymmetric=False):
        if axisymmetric:
            raise NotImplementedError("axisymmetric domains are outside this path")
        self.fes = fes
        self.integration_rule = integration_rule


class DataCache:
    """Constant coefficient (scalar, 3x3 or 6x6), src/DataCacheModule.jl:141-153."""

    def __init__(self, data):
        self.data = np.asarray(data, dtype=np.float64)

    def size(self):
        return self.data.shape


class DeforModelRed3D:
    nstressstrain = 6


class FEMMBase:
    """Finite element machine.  Keeps the device mesh of the last (geometry,
    field) it assembled, so repeated bilforms reuse the resident pattern."""

    def __init__(self, integdomain, mcsys=None, ctx=None):
        if mcsys is not None:
            raise NotImplementedError("only the identity material frame CSys(3) is on this path")
        self.integdomain = integdomain
        self.ctx = ctx
        self._mesh = None
        self._key = None

    def finite_elements(self):
        return self.integdomain.fes

    def _device_mesh(self, geom, u, nodemajor=False):
        ctx = self.ctx or default_context()
        fes = self.integdomain.fes
        if geom.values.shape[1] != 3:
            raise ValueError("geometry must be three-dimensional")
        key = (id(fes.conn), fes.conn.shape, geom.values.shape)
        if self._mesh is None or self._key != key or not np.array_equal(self._xyz, geom.values) or \
                not np.array_equal(self._conn, fes.conn):
            if self._mesh is not None and self._key == key and np.array_equal(self._conn, fes.conn):
                self._mesh.set_xyz(geom.values)  # same mesh, moved nodes: keep the pattern
            else:
                self._mesh = DeviceMesh(ctx, fes.ETYPE, fes.conn, geom.values)
                self._conn = fes.conn.copy()
                self._dofs = None
            self._xyz = geom.values.copy()
            self._key = key
        ndn = u.ndofs()
        dofs = np.arange(1, u.dofnums.size + 1, dtype=np.int64).reshape(u.dofnums.shape) if nodemajor else u.dofnums
        if self._dofs is None or not np.array_equal(self._dofs, dofs):
            if np.array_equal(dofs.ravel(), np.arange(1, dofs.size + 1)):
                self._mesh.set_dofs(ndn)
            else:
                self._mesh.set_dofs(ndn, dofs, nalldofs(u))
            self._dofs = dofs.copy()
        return self._mesh


# ------------------------------------------------------------------ assemblers
class _DeviceCOO:
    """The buffered-triples protocol shared by the sparse-matrix assemblers
    (feasm_coo of the given kind): startassembly / assemble / makematrix
    follow the reference methods of that assembler, and makematrix builds the
    CSC on the device."""

    KIND = L.ASM_SPARSE

    def _open(self, ctx, nomatrixresult):
        self.ctx = ctx or default_context()
        h = ctypes.c_void_p()
        check(lib().feasm_coo_create_kind(self.ctx.h, self.KIND, 1 if nomatrixresult else 0, ctypes.byref(h)),
              self.ctx.h)
        self.h = h
        self._nomatrixresult = bool(nomatrixresult)
        self._dims = None

    def close(self):
        if getattr(self, "h", None):
            lib().feasm_coo_destroy(self.h)
            self.h = None

    __del__ = close

    def eltype(self):
        return np.float64


class SysmatAssemblerSparse(_DeviceCOO):
    """SysmatAssemblerSparse(z=0.0, nomatrixresult=False) on the device.

    The generic protocol (startassembly / assemble / makematrix) buffers the
    triples and builds the CSC with the device sparse().  Passed to a bilform_*
    function it instead runs the GPU element loop (then ``precision`` selects
    FP64 or the mixed FP32/FP64 arithmetic)."""

    KIND = L.ASM_SPARSE

    def __init__(self, z=0.0, nomatrixresult=False, ctx=None, precision=L.FP64):
        if not isinstance(z, float):
            raise TypeError("only Float64 assemblers are supported")
        self._open(ctx, nomatrixresult)
        self.precision = precision


class SysmatAssemblerSparseSymm(_DeviceCOO):
    """SysmatAssemblerSparseSymm(z=0.0) (src/AssemblyModule.jl:342-583), the
    default assembler of the convenience bilform_* methods.  The generic
    protocol keeps each element matrix's lower triangle (local numbering) and
    makematrix forms S + transpose(S) with the diagonal halved on the device.
    Passed to a bilform_* function it runs the GPU element loop and returns
    the full symmetric matrix without the entries that sum to exactly zero,
    the pattern the reference's S + transpose(S) keeps."""

    KIND = L.ASM_SYMM

    def __init__(self, z=0.0, nomatrixresult=False, ctx=None, precision=L.FP64):
        if not isinstance(z, float):
            raise TypeError("only Float64 assemblers are supported")
        self._open(ctx, nomatrixresult)
        self.precision = precision


class SysmatAssemblerSparseDiag(_DeviceCOO):
    """SysmatAssemblerSparseDiag(z=0.0) (src/AssemblyModule.jl:599-794): the
    diagonal entries of square element matrices (off-diagonal ones ignored)."""

    KIND = L.ASM_DIAG

    def __init__(self, z=0.0, nomatrixresult=False, ctx=None):
        if not isinstance(z, float):
            raise TypeError("only Float64 assemblers are supported")
        self._open(ctx, nomatrixresult)


class SysmatAssemblerSparseHRZLumpingSymm(_DeviceCOO):
    """SysmatAssemblerSparseHRZLumpingSymm(z=0.0) (src/AssemblyModule.jl:943-1141).
    The generic protocol keeps each element matrix's diagonal scaled by its
    total over its diagonal total; passed to bilform_dot it runs the GPU
    element loop and returns the HRZ-lumped diagonal matrix (makematrix!)."""

    KIND = L.ASM_HRZ

    def __init__(self, z=0.0, nomatrixresult=False, ctx=None):
        if not isinstance(z, float):
            raise TypeError("only Float64 assemblers are supported")
        self._open(ctx, nomatrixresult)


class SysmatAssemblerFFBlock:
    """SysmatAssemblerFFBlock(row_nfreedofs, col_nfreedofs=row_nfreedofs)
    (src/AssemblyModule.jl:1149-1231): delegates startassembly / assemble to
    a wrapped SysmatAssemblerSparse and makematrix returns the leading
    free-free block matrix_blocked_ff of its matrix, extracted on the device.
    Passed to a bilform_* function it runs the GPU element loop and extracts
    the block."""

    def __init__(self, row_nfreedofs, col_nfreedofs=None, ctx=None, precision=L.FP64):
        self.row_nfree = int(row_nfreedofs)
        self.col_nfree = int(row_nfreedofs if col_nfreedofs is None else col_nfreedofs)
        self.nfree = self.row_nfree
        self.ctx = ctx or default_context()
        self.precision = precision
        self._a = SysmatAssemblerSparse(0.0, ctx=self.ctx, precision=precision)

    def eltype(self):
        return np.float64


class SysvecAssembler:
    """SysvecAssembler(z=0.0) (src/AssemblyModule.jl:853-917).

    The generic protocol keeps the reference's dense buffer semantics
    (startassembly zeroes it, assemble adds an element vector with the same
    bounds errors, makevector returns a copy).  Passed to linform_dot /
    distribloads it runs the GPU element loop instead."""

    def __init__(self, z=0.0, ctx=None):
        if not isinstance(z, float):
            raise TypeError("only Float64 assemblers are supported")
        self.ctx = ctx or default_context()
        self._buffer = np.zeros(1)
        self._row_nalldofs = 1

    def eltype(self):
        return np.float64


def setnomatrixresult(a, flag):
    if isinstance(a, SysmatAssemblerFFBlock):
        return setnomatrixresult(a._a, flag)
    a._nomatrixresult = bool(flag)
    check(lib().feasm_coo_set_nomatrixresult(a.h, 1 if flag else 0), a.ctx.h)
    return a


def expectedntriples(a, elem_mat_nrows, elem_mat_ncols, n_elem_mats):
    """expectedntriples (src/AssemblyModule.jl:54-59, :357-362, :617-622, :961-966)."""
    if isinstance(a, SysmatAssemblerSparseSymm):
        return (elem_mat_nrows * elem_mat_ncols + elem_mat_nrows) // 2 * n_elem_mats
    if isinstance(a, (SysmatAssemblerSparseDiag, SysmatAssemblerSparseHRZLumpingSymm)):
        return max(elem_mat_nrows, elem_mat_ncols) * n_elem_mats
    return elem_mat_nrows * elem_mat_ncols * n_elem_mats


def startassembly(a, elem_mat_nrows, elem_mat_ncols=None, n_elem_mats=None, row_nalldofs=None,
                  col_nalldofs=None, force_init=False):
    if isinstance(a, SysvecAssembler):  # startassembly!(self::SysvecAssembler, row_nalldofs)
        a._row_nalldofs = int(elem_mat_nrows)
        a._buffer = np.zeros(a._row_nalldofs)
        return a
    if isinstance(a, SysmatAssemblerFFBlock):  # delegated (:1183-1205)
        startassembly(a._a, elem_mat_nrows, elem_mat_ncols, n_elem_mats, row_nalldofs, col_nalldofs, force_init)
        return a
    check(lib().feasm_coo_start(a.h, elem_mat_nrows, elem_mat_ncols, n_elem_mats, row_nalldofs, col_nalldofs),
          a.ctx.h)
    if a._dims is None:
        a._dims = (int(row_nalldofs), int(col_nalldofs))
    return a


def assemble(a, mat, dofnums_row, dofnums_col=None):
    if isinstance(a, SysvecAssembler):  # assemble!(self::SysvecAssembler, vec, dofnums)
        vec = np.ravel(np.asarray(mat, dtype=np.float64))
        for i, gi in enumerate(np.ravel(dofnums_row)):
            gi = int(gi)
            if gi < 1:
                raise ValueError("Row degree of freedom < 1")
            if gi > a._row_nalldofs:
                raise ValueError("Row degree of freedom > size")
            a._buffer[gi - 1] += vec[i]
        return a
    if isinstance(a, SysmatAssemblerFFBlock):
        assemble(a._a, mat, dofnums_row, dofnums_col)
        return a
    m = np.asfortranarray(np.asarray(mat, dtype=np.float64))
    if m.ndim == 1:
        m = m.reshape(-1, 1)
    r = np.ascontiguousarray(np.ravel(dofnums_row), dtype=np.int64)
    c = np.ascontiguousarray(np.ravel(dofnums_col if dofnums_col is not None else dofnums_row), dtype=np.int64)
    check(lib().feasm_coo_assemble(a.h, m.ctypes.data_as(ctypes.c_void_p), m.shape[0], m.shape[1], ptr(r), len(r),
                                   ptr(c), len(c)), a.ctx.h)
    return a


def makevector(a):
    """makevector!(::SysvecAssembler): a copy of the buffer."""
    return a._buffer.copy()


def makematrix(a, return_device=False):
    """makematrix! of the assembler: a scipy CSC matrix (or the DeviceMatrix)."""
    if isinstance(a, SysmatAssemblerFFBlock):
        h = ctypes.c_void_p()
        check(lib().feasm_coo_makematrix(a._a.h, ctypes.byref(h)), a.ctx.h)
        if not h.value:  # nomatrixresult: the reference's spzeros, then its block
            m, n = a._a._dims if a._a._dims else (0, 0)
            if a.row_nfree > m:
                raise ValueError("The ff block has too many rows")
            if a.col_nfree > n:
                raise ValueError("The ff block has too many columns")
            return sp.csc_matrix((a.row_nfree, a.col_nfree))
        full = DeviceMatrix(a.ctx, h)
        h2 = ctypes.c_void_p()
        check(lib().feasm_matrix_ff_block(full.h, a.row_nfree, a.col_nfree, ctypes.byref(h2)), a.ctx.h)
        full.close()
        blk = DeviceMatrix(a.ctx, h2)
        return blk if return_device else blk.to_scipy()
    h = ctypes.c_void_p()
    check(lib().feasm_coo_makematrix(a.h, ctypes.byref(h)), a.ctx.h)
    if not h.value:
        m, n = a._dims if a._dims else (0, 0)
        return sp.csc_matrix((m, n))
    M = DeviceMatrix(a.ctx, h)
    return M if return_device else M.to_scipy()


# ------------------------------------------------------------------ bilinear forms
def _rule(femm):
    r = femm.integdomain.integration_rule
    fes = femm.integdomain.fes
    if isinstance(fes, FESetH8) and not isinstance(r, GaussRule):
        raise ValueError("H8 elements take a Gauss rule")
    if isinstance(fes, FESetT10) and not isinstance(r, TetRule):
        raise ValueError("T10 elements take a tetrahedral rule")
    return r.kind, r.order


def _run(femm, assembler, geom, u, form, coeff, return_device=False):
    if not isinstance(assembler, (SysmatAssemblerSparse, SysmatAssemblerSparseSymm, SysmatAssemblerFFBlock)):
        raise TypeError("the GPU bilinear forms take a SysmatAssemblerSparse, SysmatAssemblerSparseSymm or "
                        "SysmatAssemblerFFBlock")
    kind, order = _rule(femm)
    if isinstance(assembler, SysmatAssemblerSparseSymm) and assembler._nomatrixresult:
        raise NotImplementedError("nomatrixresult is not on the symmetric GPU element loop")
    if isinstance(assembler, SysmatAssemblerFFBlock):
        if assembler.col_nfree != assembler.row_nfree:
            raise NotImplementedError("the GPU element loop extracts square free-free blocks")
        if assembler.nfree > nalldofs(u):
            raise ValueError("The ff block has too many rows")
        if u._entperm_default and assembler.nfree == nfreedofs(u):
            # free-then-fixed in entity order keeps the free DOFs in node-major
            # order: assemble node-major (resident path), keep the free ones
            mesh = femm._device_mesh(geom, u, nodemajor=True)
            full = mesh.bilform(form, kind, order, coeff, precision=assembler.precision)
            keep = ~u.fixed.ravel()
        else:  # the field's own numbering (COO path), its leading block
            mesh = femm._device_mesh(geom, u)
            full = mesh.bilform(form, kind, order, coeff, precision=assembler.precision)
            keep = np.zeros(full.dims[1], bool)
            keep[: assembler.nfree] = True
        mat = full.free_block(keep)
        full.close()
        return mat if return_device else mat.to_scipy()
    mesh = femm._device_mesh(geom, u)
    mat = mesh.bilform(form, kind, order, coeff, precision=assembler.precision)
    if isinstance(assembler, SysmatAssemblerSparseSymm):
        full, mat = mat, mat.dropzeros()
        full.close()
    return mat if return_device else mat.to_scipy()


def bilform_diffusion(femm, assembler, geom, u, cf, return_device=False):
    """Diffusion stiffness: scalar DataCache -> isotropic path, 3x3 -> general."""
    if cf.data.ndim == 0 or cf.data.size == 1:
        return _run(femm, assembler, geom, u, L.DIFFUSION_ISO, [float(cf.data.ravel()[0])], return_device)
    return _run(femm, assembler, geom, u, L.DIFFUSION_MAT, cf.data.reshape(3, 3), return_device)


def bilform_lin_elastic(femm, assembler, geom, u, mr, cf, return_device=False):
    if mr is not DeforModelRed3D:
        raise NotImplementedError("only DeforModelRed3D is on this path")
    return _run(femm, assembler, geom, u, L.LIN_ELASTIC, cf.data.reshape(6, 6), return_device)


def bilform_dot(femm, assembler, geom, u, cf, m=3, return_device=False):
    if m != 3:
        raise ValueError("That is the only acceptable option here.")
    ndn = u.ndofs()
    c = np.asarray(cf.data, dtype=np.float64).reshape(ndn, ndn)
    if isinstance(assembler, SysmatAssemblerSparseHRZLumpingSymm):
        mesh = femm._device_mesh(geom, u)
        kind, order = _rule(femm)
        mat = mesh.bilform_dot_lumped(kind, order, c)
        return mat if return_device else mat.to_scipy()
    return _run(femm, assembler, geom, u, L.DOT, c, return_device)


# ------------------------------------------------------------------ linear forms
class ForceIntensity:
    """ForceIntensity with a constant force vector (its DataCache)."""

    def __init__(self, force):
        self._cache = DataCache(np.asarray(force, dtype=np.float64))


def linform_dot(femm, assembler, geom, P, f, m=3):
    """linform_dot(self, assembler, geom, P, f, m) (src/FEMMBaseModule.jl:1207-1241)
    with a constant DataCache force on a 3-manifold: the GPU element loop plus
    makevector!.  Returns the dense load vector (nalldofs entries)."""
    if not isinstance(assembler, SysvecAssembler):
        raise TypeError("the GPU linear form takes a SysvecAssembler")
    if m != 3:
        raise NotImplementedError("only body loads (m = 3) of solid elements are on this path")
    mesh = femm._device_mesh(geom, P)
    kind, order = _rule(femm)
    F = mesh.linform_dot(kind, order, np.ravel(f.data), nalldofs(P))
    assembler._row_nalldofs = nalldofs(P)
    assembler._buffer = F.copy()
    return F


def distribloads(femm, assembler, geom, P, fi, m=3):
    """distribloads (src/FEMMBaseModule.jl:1262-1297): linform_dot with the
    force intensity's cache."""
    return linform_dot(femm, assembler, geom, P, fi._cache, m)
