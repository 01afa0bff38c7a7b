"""feasm -- MI355X finite-element assembly (FinEtools.jl's element-matrix ->
SysmatAssemblerSparse -> CSC path) on hand-written gfx950 HIP kernels.

The compute lives in libfeasm.so (C ABI, include/feasm.h); this package is the
host-side mirror of the reference's interface over that ABI.
"""
from ._lib import (DIFFUSION_ISO, DIFFUSION_MAT, DOT, FP64, H8, LIN_ELASTIC, MIXED, RULE_GAUSS, RULE_TET, T10,
                   FeasmError, LIB_PATH, lib)
from .core import Batch, Context, DeviceMatrix, Server, DeviceMesh, default_context, gen_h8block, gen_perturb, gen_t10block

__all__ = [
    "DIFFUSION_ISO", "DIFFUSION_MAT", "DOT", "FP64", "H8", "LIN_ELASTIC", "MIXED", "RULE_GAUSS", "RULE_TET", "T10",
    "FeasmError", "LIB_PATH", "lib", "Batch", "Server", "Context", "DeviceMatrix", "DeviceMesh", "default_context", "gen_h8block",
    "gen_perturb", "gen_t10block",
]
