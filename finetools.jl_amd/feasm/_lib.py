"""ctypes binding of libfeasm.so (the C ABI declared in include/feasm.h).

The library is the only compute path: if it is missing this module raises,
there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfeasm.so")

OK, EINVAL, EDOF, ESIZE, EHIP, ENOMEM, ESTATE, EUNSUPPORTED, ECANCELLED = 0, -1, -2, -3, -4, -5, -6, -7, -8
H8, T10 = 0, 1
DIFFUSION_ISO, DIFFUSION_MAT, LIN_ELASTIC, DOT = 0, 1, 2, 3
RULE_GAUSS, RULE_TET = 0, 1
FP64, MIXED = 0, 1
ASM_SPARSE, ASM_SYMM, ASM_DIAG, ASM_HRZ = 0, 1, 2, 3

P = ctypes.c_void_p
I32, I64, U64, DBL = ctypes.c_int, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double
PP = ctypes.POINTER(ctypes.c_void_p)
PI64 = ctypes.POINTER(ctypes.c_int64)

# name -> (restype, argtypes); every symbol include/feasm.h declares
SIGNATURES = {
    "feasm_ctx_create": (I32, [I32, PP]),
    "feasm_ctx_destroy": (I32, [P]),
    "feasm_last_error": (ctypes.c_char_p, [P]),
    "feasm_ctx_stream": (P, [P]),
    "feasm_ctx_synchronize": (I32, [P]),
    "feasm_mem_info": (I32, [P, PI64, PI64]),
    "feasm_pool_trim": (I32, [I32, PI64]),
    "feasm_ctx_set_timing": (I32, [P, I32]),
    "feasm_timings": (I32, [P, P, I32, ctypes.POINTER(ctypes.c_int)]),
    "feasm_timings_reset": (I32, [P]),
    "feasm_copy_calibration": (I32, [P, I64, I32, ctypes.POINTER(ctypes.c_double)]),
    "feasm_probe_gram": (I32, [P, I32, I64, I32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                               P, I64]),
    "feasm_mesh_create": (I32, [P, I32, I64, I32, P, I64, P, PP]),
    "feasm_mesh_create_device": (I32, [P, I32, I64, I32, P, I64, P, PP]),
    "feasm_mesh_destroy": (I32, [P]),
    "feasm_mesh_set_xyz": (I32, [P, P]),
    "feasm_dofs_set": (I32, [P, I32, P, I64]),
    "feasm_dofs_set_nodemajor": (I32, [P, I32]),
    "feasm_bilform": (I32, [P, I32, I32, I32, P, I32, I32, PP]),
    "feasm_bilform_into": (I32, [P, I32, I32, I32, P, I32, I32, P]),
    "feasm_bilform_coo": (I32, [P, I32, I32, I32, P, I32, I32, PP]),
    "feasm_linform_dot": (I32, [P, I32, I32, P, I32, P]),
    "feasm_bilform_dot_lumped": (I32, [P, I32, I32, P, I32, PP]),
    "feasm_matrix_dropzeros": (I32, [P, PP]),
    "feasm_matrix_free_block": (I32, [P, P, PP]),
    "feasm_linform_dot_device": (I32, [P, I32, I32, P, I32, P]),
    "feasm_pattern_build": (I32, [P]),
    "feasm_mesh_set_numeric_path": (I32, [P, I32]),
    "feasm_mesh_tile_info": (I32, [P, PI64, PI64]),
    "feasm_tiles_compare": (I32, [P, I32, I32, PI64, PI64, ctypes.POINTER(DBL)]),
    "feasm_mesh_set_tile_params": (I32, [P, I32, I32, I32]),
    "feasm_pattern_info": (I32, [P, PI64, PI64]),
    "feasm_matrix_dims": (I32, [P, PI64, PI64, PI64]),
    "feasm_matrix_copy_csc": (I32, [P, I32, P, P, P]),
    "feasm_matrix_device_csc": (I32, [P, PP, PP, PP]),
    "feasm_matrix_destroy": (I32, [P]),
    "feasm_matrix_checksum": (I32, [P, ctypes.POINTER(ctypes.c_uint64)]),
    "feasm_coo_create": (I32, [P, I32, PP]),
    "feasm_coo_create_kind": (I32, [P, I32, I32, PP]),
    "feasm_matrix_ff_block": (I32, [P, I64, I64, PP]),
    "feasm_coo_start": (I32, [P, I64, I64, I64, I64, I64]),
    "feasm_coo_assemble": (I32, [P, P, I64, I64, P, I64, P, I64]),
    "feasm_coo_assemble_many": (I32, [P, I64, P, P, I64, P, I64]),
    "feasm_coo_set_nomatrixresult": (I32, [P, I32]),
    "feasm_coo_makematrix": (I32, [P, PP]),
    "feasm_coo_ntriples": (I64, [P]),
    "feasm_coo_destroy": (I32, [P]),
    "feasm_sparse_device": (I32, [P, I64, P, P, P, I64, I64, PP]),
    "feasm_sparse": (I32, [P, I64, P, P, P, I64, I64, PP]),
    "feasm_sparse_device_noalloc": (I32, [P, I64, P, P, P, I64, I64, P, P, P, PI64]),
    "feasm_batch_create": (I32, [P, PP]),
    "feasm_batch_add": (I32, [P, P, I32, I32, I32, P, I32, I32]),
    "feasm_batch_prepare": (I32, [P, I32, I32, I32, I32]),
    "feasm_batch_run": (I32, [P]),
    "feasm_batch_last_ms": (I32, [P, ctypes.POINTER(ctypes.c_double)]),
    "feasm_batch_matrix": (I32, [P, I64, PP]),
    "feasm_batch_stats": (I32, [P, PI64, PI64, PI64, PI64]),
    "feasm_batch_autotune": (I32, [P, ctypes.c_char_p, ctypes.c_char_p, I32, ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_double)]),
    "feasm_batch_destroy": (I32, [P]),
    "feasm_server_create": (I32, [I32, I32, DBL, PP]),
    "feasm_server_submit_mesh": (I32, [P, I32, I64, I32, P, I64, P, I32, I32, I32, I32, P, I32, I32, I32, PI64, PI64]),
    "feasm_server_submit_mesh_borrowed": (I32, [P, I32, I64, I32, P, I64, P, I32, I32, I32, I32, P, I32, I32, I32, PI64,
                                                PI64]),
    "feasm_server_submit_reassembly": (I32, [P, I64, I32, I32, I32, P, I32, I32, PI64]),
    "feasm_server_wait": (I32, [P, I64, DBL, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_double)]),
    "feasm_server_cancel": (I32, [P, I64]),
    "feasm_server_result": (I32, [P, I64, PP]),
    "feasm_server_release": (I32, [P, I64]),
    "feasm_server_stats": (I32, [P, PI64, PI64, PI64, PI64, PI64, PI64]),
    "feasm_server_last_error": (ctypes.c_char_p, [P]),
    "feasm_server_destroy": (I32, [P]),
    "feasm_gen_h8block": (I32, [DBL, DBL, DBL, I64, I64, I64, PI64, PI64, P, P]),
    "feasm_gen_t10block": (I32, [DBL, DBL, DBL, I64, I64, I64, DBL, DBL, U64, PI64, PI64, P, P]),
    "feasm_gen_perturb": (I32, [I64, P, P, P, DBL, U64]),
}

_lib = None


class FeasmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libfeasm.so is not built (%s); run `python finetools.jl_amd/build.py` "
                "or __graft_entry__.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, ctx=None):
    if rc != OK:
        msg = lib().feasm_last_error(ctx)
        raise FeasmError(rc, msg.decode() if msg else "feasm error %d" % rc)
    return rc


def ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)
