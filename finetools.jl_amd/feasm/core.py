"""Thin object layer over the C ABI: device context, resident meshes, matrices.

These objects own library handles (device memory) and release them when
garbage-collected or closed.
"""
import ctypes

import numpy as np
import scipy.sparse as sp

from . import _lib as L
from ._lib import check, lib, ptr

_default_ctx = None


class Context:
    """One HIP device + one stream (feasm_ctx)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(lib().feasm_ctx_create(device, ctypes.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            lib().feasm_ctx_destroy(self.h)
            self.h = None

    __del__ = close

    @property
    def stream(self):
        return lib().feasm_ctx_stream(self.h)

    def synchronize(self):
        check(lib().feasm_ctx_synchronize(self.h), self.h)

    def mem_info(self):
        f, t = ctypes.c_int64(), ctypes.c_int64()
        check(lib().feasm_mem_info(self.h, ctypes.byref(f), ctypes.byref(t)), self.h)
        return f.value, t.value

    def set_timing(self, slots=64):
        """Record HIP events around the kernels of the next `slots` numeric
        assemblies (0 = off)."""
        check(lib().feasm_ctx_set_timing(self.h, int(slots)), self.h)

    def timings(self, max_calls=4096):
        """(n, 3) array of [element_ms, gather_ms, sum] per recorded call."""
        out = np.zeros(3 * max_calls)
        n = ctypes.c_int()
        check(lib().feasm_timings(self.h, ptr(out), max_calls, ctypes.byref(n)), self.h)
        return out[:3 * n.value].reshape(n.value, 3)

    def reset_timings(self):
        check(lib().feasm_timings_reset(self.h), self.h)


def pool_trim(device=0):
    """Free the device memory the library keeps for reuse; the bytes freed."""
    f = ctypes.c_int64()
    check(lib().feasm_pool_trim(device, ctypes.byref(f)))
    return f.value


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class DeviceMatrix:
    """A CSC matrix held in HBM (feasm_matrix)."""

    def __init__(self, ctx, handle, owner=None):
        self.ctx = ctx
        self.h = handle
        self._owner = owner  # set when another handle owns the matrix (batch results)

    def close(self):
        if getattr(self, "h", None):
            if self._owner is None:
                lib().feasm_matrix_destroy(self.h)
            self.h = None

    __del__ = close

    @property
    def dims(self):
        m, n, z = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(lib().feasm_matrix_dims(self.h, ctypes.byref(m), ctypes.byref(n), ctypes.byref(z)))
        return m.value, n.value, z.value

    @property
    def nnz(self):
        return self.dims[2]

    def dropzeros(self):
        """A new DeviceMatrix without the stored exact zeros (SparseArrays.dropzeros)."""
        h = ctypes.c_void_p()
        check(lib().feasm_matrix_dropzeros(self.h, ctypes.byref(h)), self.ctx.h)
        return DeviceMatrix(self.ctx, h)

    def free_block(self, is_free):
        """The principal submatrix on the DOFs flagged free (one flag per column,
        0-based), renumbered consecutively: matrix_blocked_ff of the
        free-then-fixed numbering."""
        f = np.ascontiguousarray(np.asarray(is_free, dtype=np.uint8).ravel())
        if f.size != self.dims[1]:
            raise ValueError("one flag per column")
        h = ctypes.c_void_p()
        check(lib().feasm_matrix_free_block(self.h, ptr(f), ctypes.byref(h)), self.ctx.h)
        return DeviceMatrix(self.ctx, h)

    def csc_arrays(self, base=0):
        m, n, nnz = self.dims
        colptr = np.empty(n + 1, np.int64)
        rowval = np.empty(nnz, np.int64)
        nzval = np.empty(nnz, np.float64)
        check(lib().feasm_matrix_copy_csc(self.h, base, ptr(colptr), ptr(rowval), ptr(nzval)),
              self.ctx.h if self.ctx is not None else None)
        return colptr, rowval, nzval

    def to_scipy(self):
        m, n, _ = self.dims
        colptr, rowval, nzval = self.csc_arrays(0)
        return sp.csc_matrix((nzval, rowval, colptr), shape=(m, n))

    def checksum(self):
        """64-bit digest of colptr, rowval and the value bits (computed on the device)."""
        d = ctypes.c_uint64()
        check(lib().feasm_matrix_checksum(self.h, ctypes.byref(d)), self.ctx.h if self.ctx is not None else None)
        return d.value

    def device_ptrs(self):
        c, r, v = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().feasm_matrix_device_csc(self.h, ctypes.byref(c), ctypes.byref(r), ctypes.byref(v)))
        return c.value, r.value, v.value


class DeviceMesh:
    """Connectivity + coordinates + DOF numbering resident in HBM (feasm_mesh).

    conn: (nelem, npe) 1-based node numbers; xyz: (nnodes, 3)."""

    def __init__(self, ctx, etype, conn, xyz):
        conn = np.ascontiguousarray(conn, dtype=np.int64)
        xyz = np.asarray(xyz, dtype=np.float64)
        self.ctx = ctx
        self.etype = etype
        self.nelem, self.npe = conn.shape
        self.nnodes = xyz.shape[0]
        xyz_cm = np.asfortranarray(xyz)  # Julia's column-major nnodes x 3
        h = ctypes.c_void_p()
        check(lib().feasm_mesh_create(ctx.h, etype, self.nelem, self.npe, ptr(conn), self.nnodes,
                                      xyz_cm.ctypes.data_as(ctypes.c_void_p), ctypes.byref(h)), ctx.h)
        self.h = h
        self.ndn = 0

    def close(self):
        if getattr(self, "h", None):
            lib().feasm_mesh_destroy(self.h)
            self.h = None

    __del__ = close

    def set_xyz(self, xyz):
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        check(lib().feasm_mesh_set_xyz(self.h, ptr(x)), self.ctx.h)

    def set_dofs(self, ndn, dofnums=None, nalldofs=None):
        if dofnums is None:
            check(lib().feasm_dofs_set_nodemajor(self.h, ndn), self.ctx.h)
        else:
            d = np.asfortranarray(np.asarray(dofnums, dtype=np.int64).reshape(self.nnodes, ndn))
            check(lib().feasm_dofs_set(self.h, ndn, d.ctypes.data_as(ctypes.c_void_p), int(nalldofs)),
                  self.ctx.h)
        self.ndn = ndn

    def pattern_info(self):
        nnz, nb = ctypes.c_int64(), ctypes.c_int64()
        check(lib().feasm_pattern_info(self.h, ctypes.byref(nnz), ctypes.byref(nb)), self.ctx.h)
        return nnz.value, nb.value

    def set_numeric_path(self, path):
        """'auto', 'staged' or 'fused'."""
        check(lib().feasm_mesh_set_numeric_path(self.h, {"auto": 0, "staged": 1, "fused": 2}[path]), self.ctx.h)

    def set_tile_params(self, max_nodes=0, budget_kib=0, threads=0):
        check(lib().feasm_mesh_set_tile_params(self.h, max_nodes, budget_kib, threads), self.ctx.h)

    def tile_info(self):
        t, h = ctypes.c_int64(), ctypes.c_int64()
        check(lib().feasm_mesh_tile_info(self.h, ctypes.byref(t), ctypes.byref(h)), self.ctx.h)
        return t.value, h.value

    def build_pattern(self):
        check(lib().feasm_pattern_build(self.h), self.ctx.h)

    def tiles_compare(self, form, precision=L.FP64):
        """Builds the fused kernel's tiles with the device and the host
        builders and compares them byte for byte: (ntiles, first differing
        tile or -1, device build ms)."""
        nt, diff, ms = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_double()
        check(lib().feasm_tiles_compare(self.h, form, precision, ctypes.byref(nt), ctypes.byref(diff),
                                        ctypes.byref(ms)), self.ctx.h)
        return nt.value, diff.value, ms.value

    def bilform(self, form, rule, order, coeff, precision=L.FP64, path="auto"):
        c = np.ascontiguousarray(np.atleast_1d(coeff), dtype=np.float64).ravel()
        h = ctypes.c_void_p()
        fn = lib().feasm_bilform_coo if path == "coo" else lib().feasm_bilform
        check(fn(self.h, form, rule, order, ptr(c), len(c), precision, ctypes.byref(h)), self.ctx.h)
        return DeviceMatrix(self.ctx, h)

    def bilform_into(self, mat, form, rule, order, coeff, precision=L.FP64):
        c = np.ascontiguousarray(np.atleast_1d(coeff), dtype=np.float64).ravel()
        check(lib().feasm_bilform_into(self.h, form, rule, order, ptr(c), len(c), precision, mat.h), self.ctx.h)
        return mat

    def bilform_dot_lumped(self, rule, order, coeff):
        """HRZ-lumped mass (bilform_dot into SysmatAssemblerSparseHRZLumpingSymm,
        makematrix!): a diagonal DeviceMatrix."""
        c = np.ascontiguousarray(np.atleast_1d(coeff), dtype=np.float64).ravel()
        h = ctypes.c_void_p()
        check(lib().feasm_bilform_dot_lumped(self.h, rule, order, ptr(c), len(c), ctypes.byref(h)), self.ctx.h)
        return DeviceMatrix(self.ctx, h)

    def linform_dot(self, rule, order, force, nalldofs):
        """Load vector of a constant force (linform_dot, m = 3, + makevector!):
        a host array of nalldofs entries, entry d-1 for DOF d."""
        f = np.ascontiguousarray(np.atleast_1d(force), dtype=np.float64).ravel()
        F = np.zeros(int(nalldofs))
        check(lib().feasm_linform_dot(self.h, rule, order, ptr(f), len(f), ptr(F)), self.ctx.h)
        return F


class Batch:
    """Many meshes assembled by one persistent launch (feasm_batch)."""

    def __init__(self, ctx):
        h = ctypes.c_void_p()
        check(lib().feasm_batch_create(ctx.h, ctypes.byref(h)), ctx.h)
        self.ctx = ctx
        self.h = h
        self.meshes = []

    def close(self):
        if getattr(self, "h", None):
            lib().feasm_batch_destroy(self.h)
            self.h = None

    __del__ = close

    def add(self, mesh, form, rule, order, coeff, precision=L.FP64):
        c = np.ascontiguousarray(np.atleast_1d(coeff), dtype=np.float64).ravel()
        check(lib().feasm_batch_add(self.h, mesh.h, form, rule, order, ptr(c), len(c), precision), self.ctx.h)
        self.meshes.append(mesh)
        return len(self.meshes) - 1

    def prepare(self, use_graph=True, max_nodes=0, budget_kib=0, threads=0):
        check(lib().feasm_batch_prepare(self.h, 1 if use_graph else 0, max_nodes, budget_kib, threads), self.ctx.h)

    def run(self):
        check(lib().feasm_batch_run(self.h), self.ctx.h)

    def last_ms(self):
        v = ctypes.c_double()
        check(lib().feasm_batch_last_ms(self.h, ctypes.byref(v)), self.ctx.h)
        return v.value

    def matrix(self, i):
        h = ctypes.c_void_p()
        check(lib().feasm_batch_matrix(self.h, i, ctypes.byref(h)), self.ctx.h)
        return DeviceMatrix(self.ctx, h, owner=self)

    def stats(self):
        a, b, c, d = (ctypes.c_int64() for _ in range(4))
        check(lib().feasm_batch_stats(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d)),
              self.ctx.h)
        return {"meshes": a.value, "elements": b.value, "nnz": c.value, "tiles": d.value}

    def autotune(self, db_path=None, key=None, reps=3):
        n, kb, th, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        check(lib().feasm_batch_autotune(self.h, db_path.encode() if db_path else None,
                                         key.encode() if key else None, reps, ctypes.byref(n), ctypes.byref(kb),
                                         ctypes.byref(th), ctypes.byref(ms)), self.ctx.h)
        return {"max_nodes": n.value, "budget_kib": kb.value, "threads": th.value, "ms": ms.value}


class Server:
    """Continuous-batching assembly server (feasm_server)."""

    def __init__(self, device=0, workers=4, hbm_fraction=0.5):
        h = ctypes.c_void_p()
        rc = lib().feasm_server_create(device, workers, hbm_fraction, ctypes.byref(h))
        if rc != L.OK:
            raise L.FeasmError(rc, lib().feasm_server_last_error(None).decode())
        self.h = h
        self._borrowed = {}  # request id -> arrays the server reads until the request ends

    def _check(self, rc):
        if rc != L.OK:
            raise L.FeasmError(rc, lib().feasm_server_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().feasm_server_destroy(self.h)
            self.h = None
        self._borrowed = {}

    __del__ = close

    def submit_mesh(self, etype, conn, xyz, ndn, form, rule, order, coeff, precision=L.FP64, keep=False,
                    borrow=False):
        """borrow: the server reads conn and xyz in its worker, without a copy;
        the arrays are held here until the request is released."""
        conn = np.ascontiguousarray(conn, dtype=np.int64)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        c = np.ascontiguousarray(np.atleast_1d(coeff), dtype=np.float64).ravel()
        rid, mid = ctypes.c_int64(), ctypes.c_int64()
        fn = lib().feasm_server_submit_mesh_borrowed if borrow else lib().feasm_server_submit_mesh
        self._check(fn(self.h, etype, conn.shape[0], conn.shape[1], ptr(conn), xyz.shape[0], ptr(xyz), ndn, form, rule,
                       order, ptr(c), len(c), precision, 1 if keep else 0, ctypes.byref(rid), ctypes.byref(mid)))
        if borrow:
            self._borrowed[rid.value] = (conn, xyz)
        return rid.value, mid.value

    def submit_reassembly(self, mesh_id, form, rule, order, coeff, precision=L.FP64):
        c = np.ascontiguousarray(np.atleast_1d(coeff), dtype=np.float64).ravel()
        rid = ctypes.c_int64()
        self._check(lib().feasm_server_submit_reassembly(self.h, mesh_id, form, rule, order, ptr(c), len(c),
                                                         precision, ctypes.byref(rid)))
        return rid.value

    def wait(self, rid, timeout=600.0):
        st, lat, qm = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        self._check(lib().feasm_server_wait(self.h, rid, timeout, ctypes.byref(st), ctypes.byref(lat),
                                            ctypes.byref(qm)))
        return st.value, lat.value, qm.value

    def cancel(self, rid):
        """Withdraw a queued request, or have a running one reported cancelled."""
        self._check(lib().feasm_server_cancel(self.h, rid))

    def result(self, rid):
        h = ctypes.c_void_p()
        self._check(lib().feasm_server_result(self.h, rid, ctypes.byref(h)))
        return DeviceMatrix(None, h, owner=self)

    def release(self, rid):
        self._check(lib().feasm_server_release(self.h, rid))
        self._borrowed.pop(rid, None)

    def stats(self):
        v = [ctypes.c_int64() for _ in range(6)]
        self._check(lib().feasm_server_stats(self.h, *[ctypes.byref(x) for x in v]))
        return dict(zip(["queued", "running", "completed", "residents", "bytes_in_use", "budget"],
                        [x.value for x in v]))


def sparse(ctx, I, J, V, m, n):
    """Device sparse(I, J, V, m, n) of host triples with 1-based indices."""
    I = np.ascontiguousarray(I, dtype=np.int64)
    J = np.ascontiguousarray(J, dtype=np.int64)
    V = np.ascontiguousarray(V, dtype=np.float64)
    h = ctypes.c_void_p()
    check(lib().feasm_sparse(ctx.h, len(V), ptr(I), ptr(J), ptr(V), m, n, ctypes.byref(h)), ctx.h)
    return DeviceMatrix(ctx, h)


# ------------------------------------------------------------------ generators
def gen_h8block(L_, W, H, nL, nW, nH):
    ne, nn = ctypes.c_int64(), ctypes.c_int64()
    check(lib().feasm_gen_h8block(L_, W, H, nL, nW, nH, ctypes.byref(ne), ctypes.byref(nn), None, None))
    conn = np.empty((ne.value, 8), np.int64)
    xyz = np.empty((nn.value, 3))
    check(lib().feasm_gen_h8block(L_, W, H, nL, nW, nH, ctypes.byref(ne), ctypes.byref(nn), ptr(conn), ptr(xyz)))
    return conn, xyz


def gen_t10block(L_, W, H, nL, nW, nH, amp_vertex=0.0, amp_mid=0.0, seed=0):
    ne, nn = ctypes.c_int64(), ctypes.c_int64()
    check(lib().feasm_gen_t10block(L_, W, H, nL, nW, nH, amp_vertex, amp_mid, seed, ctypes.byref(ne),
                                   ctypes.byref(nn), None, None))
    conn = np.empty((ne.value, 10), np.int64)
    xyz = np.empty((nn.value, 3))
    check(lib().feasm_gen_t10block(L_, W, H, nL, nW, nH, amp_vertex, amp_mid, seed, ctypes.byref(ne),
                                   ctypes.byref(nn), ptr(conn), ptr(xyz)))
    return conn, xyz


def gen_perturb(xyz, lo, hi, amp, seed):
    x = np.ascontiguousarray(xyz, dtype=np.float64).copy()
    lo = np.ascontiguousarray(lo, dtype=np.float64)
    hi = np.ascontiguousarray(hi, dtype=np.float64)
    check(lib().feasm_gen_perturb(x.shape[0], ptr(x), ptr(lo), ptr(hi), amp, seed))
    return x
