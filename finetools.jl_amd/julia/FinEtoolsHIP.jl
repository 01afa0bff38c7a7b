"""
    FinEtoolsHIP

Julia side of the drop-in: a `SysmatAssemblerSparse`-compatible assembler
backed by libfeasm.so, plus `bilform_*` methods specialised on it so that the
whole element loop of `bilform_diffusion`, `bilform_lin_elastic` and
`bilform_dot` runs on an MI355X, with the reference's call signatures.

    using FinEtools, FinEtoolsHIP
    a = SysmatAssemblerSparseHIP(0.0)
    K = bilform_lin_elastic(femm, a, geom, u, DeforModelRed3D, DataCache(D))

Generic protocol calls (startassembly!/assemble!/makematrix!) on the same
assembler buffer the triples and build the CSC with the device sparse().
Anything the device path does not cover (non-constant DataCache, a material
frame other than the identity, other element types/rules) falls back to
FinEtools' own generic method through `invoke`.

This file is not exercised in the build container (no Julia there); the C ABI
it binds is tested from Python (tests/test_gpu_parity.py).
"""
module FinEtoolsHIP

using FinEtools
using FinEtools.AssemblyModule: AbstractSysmatAssembler, AbstractSysvecAssembler
import FinEtools.AssemblyModule: startassembly!, assemble!, makematrix!, makevector!
import FinEtools.FEMMBaseModule: bilform_diffusion, bilform_lin_elastic, bilform_dot, linform_dot
using SparseArrays: SparseMatrixCSC, spzeros

const libfeasm = get(ENV, "FEASM_LIB", joinpath(@__DIR__, "..", "feasm", "libfeasm.so"))

const FEASM_H8, FEASM_T10 = Cint(0), Cint(1)
const FORM_DIFF_ISO, FORM_DIFF_MAT, FORM_LIN_ELASTIC, FORM_DOT = Cint(0), Cint(1), Cint(2), Cint(3)
const RULE_GAUSS, RULE_TET = Cint(0), Cint(1)

function check(rc, ctx)
    rc == 0 && return
    msg = unsafe_string(ccall((:feasm_last_error, libfeasm), Cstring, (Ptr{Cvoid},), ctx))
    error(msg)  # the reference's message text, e.g. "Row degree of freedom < 1"
end

const _ctx = Ref{Ptr{Cvoid}}(C_NULL)
function context()
    if _ctx[] == C_NULL
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:feasm_ctx_create, libfeasm), Cint, (Cint, Ptr{Ptr{Cvoid}}), 0, h), C_NULL)
        _ctx[] = h[]
    end
    return _ctx[]
end

mutable struct SysmatAssemblerSparseHIP{T} <: AbstractSysmatAssembler
    _matbuffer::Vector{T}      # kept so that eltype(a) works (AssemblyModule.jl:27)
    _coo::Ptr{Cvoid}
    _nomatrixresult::Bool
    _force_init::Bool
    _row_nalldofs::Int
    _col_nalldofs::Int
    _meshes::Dict{UInt,Ptr{Cvoid}}   # device meshes keyed by (fes, geom, u) identity
end

function SysmatAssemblerSparseHIP(z::Float64 = 0.0, nomatrixresult = false)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:feasm_coo_create, libfeasm), Cint, (Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}),
                context(), nomatrixresult, h), context())
    a = SysmatAssemblerSparseHIP{Float64}(Float64[z], h[], nomatrixresult, false, 0, 0, Dict{UInt,Ptr{Cvoid}}())
    finalizer(a) do x
        ccall((:feasm_coo_destroy, libfeasm), Cint, (Ptr{Cvoid},), x._coo)
        for m in values(x._meshes)
            ccall((:feasm_mesh_destroy, libfeasm), Cint, (Ptr{Cvoid},), m)
        end
    end
    return a
end

# ---------------------------------------------------------------- generic protocol
function startassembly!(self::SysmatAssemblerSparseHIP, elem_mat_nrows::IT, elem_mat_ncols::IT,
                        n_elem_mats::IT, row_nalldofs::IT, col_nalldofs::IT; force_init = false) where {IT<:Integer}
    check(ccall((:feasm_coo_start, libfeasm), Cint, (Ptr{Cvoid}, Int64, Int64, Int64, Int64, Int64),
                self._coo, elem_mat_nrows, elem_mat_ncols, n_elem_mats, row_nalldofs, col_nalldofs), context())
    if self._row_nalldofs == 0
        self._row_nalldofs, self._col_nalldofs = row_nalldofs, col_nalldofs
    end
    self._force_init = force_init
    return self
end

function assemble!(self::SysmatAssemblerSparseHIP, mat::AbstractMatrix, dofnums_row, dofnums_col)
    m = Matrix{Float64}(mat)
    r = Vector{Int64}(vec(dofnums_row))
    c = Vector{Int64}(vec(dofnums_col))
    check(ccall((:feasm_coo_assemble, libfeasm), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Int64, Int64, Ptr{Int64}, Int64, Ptr{Int64}, Int64),
                self._coo, m, size(m, 1), size(m, 2), r, length(r), c, length(c)), context())
    return self
end

function _tocsc(h::Ptr{Cvoid})
    m, n, nnz = Ref{Int64}(0), Ref{Int64}(0), Ref{Int64}(0)
    check(ccall((:feasm_matrix_dims, libfeasm), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}, Ref{Int64}),
                h, m, n, nnz), context())
    colptr = Vector{Int64}(undef, n[] + 1)
    rowval = Vector{Int64}(undef, nnz[])
    nzval = Vector{Float64}(undef, nnz[])
    check(ccall((:feasm_matrix_copy_csc, libfeasm), Cint, (Ptr{Cvoid}, Cint, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}),
                h, 1, colptr, rowval, nzval), context())
    ccall((:feasm_matrix_destroy, libfeasm), Cint, (Ptr{Cvoid},), h)
    return SparseMatrixCSC{Float64,Int64}(m[], n[], colptr, rowval, nzval)
end

function makematrix!(self::SysmatAssemblerSparseHIP)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:feasm_coo_set_nomatrixresult, libfeasm), Cint, (Ptr{Cvoid}, Cint), self._coo,
                self._nomatrixresult), context())
    check(ccall((:feasm_coo_makematrix, libfeasm), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}), self._coo, h), context())
    h[] == C_NULL && return spzeros(self._row_nalldofs, self._col_nalldofs)
    return _tocsc(h[])
end

# ---------------------------------------------------------------- GPU bilinear forms
_etype(::FESetH8) = FEASM_H8
_etype(::FESetT10) = FEASM_T10
_etype(::Any) = nothing
_rule(r::GaussRule) = (RULE_GAUSS, r.dim)   # GaussRule stores order in `dim` (IntegRuleModule.jl:391-392)
_rule(r::TetRule) = (RULE_TET, r.npts)
_rule(::Any) = nothing

# device meshes by (element set, geometry, field numbering), shared by all
# the assemblers of this module
const _MESHES = Dict{UInt,Ptr{Cvoid}}()

function _mesh_for(femm, geom, u)
    fes = finite_elements(femm)
    key = hash((objectid(fes), objectid(geom), objectid(u), geom.values, u.dofnums))
    haskey(_MESHES, key) && return _MESHES[key]
    conn = Int64[c for e in fes.conn for c in e]                 # nelem x npe, row-major
    xyz = Matrix{Float64}(geom.values)                            # nnodes x 3, column-major
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:feasm_mesh_create, libfeasm), Cint,
                (Ptr{Cvoid}, Cint, Int64, Cint, Ptr{Int64}, Int64, Ptr{Float64}, Ptr{Ptr{Cvoid}}),
                context(), _etype(fes), count(fes), nodesperelem(fes), conn, size(xyz, 1), xyz, h), context())
    dn = Matrix{Int64}(u.dofnums)
    check(ccall((:feasm_dofs_set, libfeasm), Cint, (Ptr{Cvoid}, Cint, Ptr{Int64}, Int64),
                h[], size(dn, 2), dn, nalldofs(u)), context())
    _MESHES[key] = h[]
    return h[]
end

_device_mesh(self::SysmatAssemblerSparseHIP, femm, geom, u) = _mesh_for(femm, geom, u)

function _gpu_bilform_handle(femm, geom, u, form, coeff::Vector{Float64})
    kind, order = _rule(femm.integdomain.integration_rule)
    mesh = _mesh_for(femm, geom, u)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:feasm_bilform, libfeasm), Cint,
                (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Float64}, Cint, Cint, Ptr{Ptr{Cvoid}}),
                mesh, form, kind, order, coeff, length(coeff), 0, h), context())
    return h[]
end

_gpu_bilform(self::SysmatAssemblerSparseHIP, femm, geom, u, form, coeff::Vector{Float64}) =
    _tocsc(_gpu_bilform_handle(femm, geom, u, form, coeff))

# The device path applies to a constant DataCache (its fill function is the
# constant closure of DataCache(data), DataCacheModule.jl:76-89), the identity
# material frame (CSys(3), CSysModule.jl:154-181) and H8/T10 with a
# Gauss/Tet rule; everything else defers to FinEtools' generic methods.
_constant(cf::DataCache) = occursin("fillcache_constant", string(nameof(typeof(cf._fillcache!))))
_onpath(femm, cf) = (_etype(finite_elements(femm)) !== nothing) &&
                    (_rule(femm.integdomain.integration_rule) !== nothing) &&
                    femm.mcsys.isidentity && _constant(cf)

function bilform_diffusion(self::FEMM, assembler::SysmatAssemblerSparseHIP, geom::NodalField{FT},
                           u::NodalField{T}, cf::DC) where {FEMM<:AbstractFEMM,FT,T,DC<:DataCache}
    if !_onpath(self, cf)
        return invoke(bilform_diffusion, Tuple{AbstractFEMM,AbstractSysmatAssembler,NodalField,NodalField,DataCache},
                      self, assembler, geom, u, cf)
    end
    c = cf(zeros(1, 3), zeros(3, 3), 0, 0)
    isempty(size(c)) && return _gpu_bilform(assembler, self, geom, u, FORM_DIFF_ISO, [Float64(c)])
    return _gpu_bilform(assembler, self, geom, u, FORM_DIFF_MAT, vec(Matrix{Float64}(c)'))
end

function bilform_lin_elastic(self::FEMM, assembler::SysmatAssemblerSparseHIP, geom::NodalField{FT},
                             u::NodalField{T}, mr::Type{DeforModelRed3D}, cf::DC) where {FEMM<:AbstractFEMM,FT,T,DC<:DataCache}
    if !_onpath(self, cf)
        return invoke(bilform_lin_elastic,
                      Tuple{AbstractFEMM,AbstractSysmatAssembler,NodalField,NodalField,Type,DataCache},
                      self, assembler, geom, u, mr, cf)
    end
    D = cf(zeros(1, 3), zeros(3, 3), 0, 0)
    return _gpu_bilform(assembler, self, geom, u, FORM_LIN_ELASTIC, vec(Matrix{Float64}(D)'))
end

function bilform_dot(self::FEMM, assembler::SysmatAssemblerSparseHIP, geom::NodalField{FT}, u::NodalField{T},
                     cf::DC; m = 3) where {FEMM<:AbstractFEMM,FT,T,DC<:DataCache}
    if !_onpath(self, cf) || m != 3
        return invoke(bilform_dot, Tuple{AbstractFEMM,AbstractSysmatAssembler,NodalField,NodalField,DataCache},
                      self, assembler, geom, u, cf; m = m)
    end
    c = cf(zeros(1, 3), zeros(3, 3), 0, 0)
    return _gpu_bilform(assembler, self, geom, u, FORM_DOT, vec(Matrix{Float64}(c)'))
end

# ---------------------------------------------------------------- further assemblers
"""
    SysmatAssemblerSparseSymmHIP()

In place of `SysmatAssemblerSparseSymm` (src/AssemblyModule.jl:342-583): the
full symmetric matrix from the GPU element loop, without the entries that sum
to exactly zero (the pattern the reference's `S + transpose(S)` keeps).
"""
struct SysmatAssemblerSparseSymmHIP <: AbstractSysmatAssembler end

"""
    SysmatAssemblerSparseHRZLumpingSymmHIP()

In place of `SysmatAssemblerSparseHRZLumpingSymm` (src/AssemblyModule.jl:943-1141)
for `bilform_dot`: the HRZ-lumped diagonal mass computed on the GPU.
"""
struct SysmatAssemblerSparseHRZLumpingSymmHIP <: AbstractSysmatAssembler end

"""
    SysmatAssemblerFFBlockHIP(row_nfreedofs)

In place of `SysmatAssemblerFFBlock` (src/AssemblyModule.jl:1149-1231): the GPU
matrix in the field's numbering, then its leading free-free block extracted on
the device (`matrix_blocked_ff`, src/MatrixUtilityModule.jl:675-690).
"""
struct SysmatAssemblerFFBlockHIP <: AbstractSysmatAssembler
    nfree::Int
end

function _finish(::SysmatAssemblerSparseSymmHIP, h::Ptr{Cvoid}, u)
    h2 = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:feasm_matrix_dropzeros, libfeasm), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}), h, h2), context())
    ccall((:feasm_matrix_destroy, libfeasm), Cint, (Ptr{Cvoid},), h)
    return _tocsc(h2[])
end

function _finish(a::SysmatAssemblerFFBlockHIP, h::Ptr{Cvoid}, u)
    keep = UInt8[d <= a.nfree ? 1 : 0 for d in 1:nalldofs(u)]
    h2 = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:feasm_matrix_free_block, libfeasm), Cint, (Ptr{Cvoid}, Ptr{UInt8}, Ptr{Ptr{Cvoid}}),
                h, keep, h2), context())
    ccall((:feasm_matrix_destroy, libfeasm), Cint, (Ptr{Cvoid},), h)
    return _tocsc(h2[])
end

const _WRAPPED = Union{SysmatAssemblerSparseSymmHIP,SysmatAssemblerFFBlockHIP}

_gpu_bilform(self::_WRAPPED, femm, geom, u, form, coeff::Vector{Float64}) =
    _finish(self, _gpu_bilform_handle(femm, geom, u, form, coeff), u)

for A in (:SysmatAssemblerSparseSymmHIP, :SysmatAssemblerFFBlockHIP)
    @eval begin
        function bilform_diffusion(self::FEMM, assembler::$A, geom::NodalField{FT}, u::NodalField{T},
                                   cf::DC) where {FEMM<:AbstractFEMM,FT,T,DC<:DataCache}
            _onpath(self, cf) || error("this assembler needs a constant DataCache, CSys(3) and H8/T10")
            c = cf(zeros(1, 3), zeros(3, 3), 0, 0)
            isempty(size(c)) && return _gpu_bilform(assembler, self, geom, u, FORM_DIFF_ISO, [Float64(c)])
            return _gpu_bilform(assembler, self, geom, u, FORM_DIFF_MAT, vec(Matrix{Float64}(c)'))
        end
        function bilform_lin_elastic(self::FEMM, assembler::$A, geom::NodalField{FT}, u::NodalField{T},
                                     mr::Type{DeforModelRed3D}, cf::DC) where {FEMM<:AbstractFEMM,FT,T,DC<:DataCache}
            _onpath(self, cf) || error("this assembler needs a constant DataCache, CSys(3) and H8/T10")
            D = cf(zeros(1, 3), zeros(3, 3), 0, 0)
            return _gpu_bilform(assembler, self, geom, u, FORM_LIN_ELASTIC, vec(Matrix{Float64}(D)'))
        end
        function bilform_dot(self::FEMM, assembler::$A, geom::NodalField{FT}, u::NodalField{T}, cf::DC;
                             m = 3) where {FEMM<:AbstractFEMM,FT,T,DC<:DataCache}
            (_onpath(self, cf) && m == 3) || error("this assembler needs a constant DataCache, CSys(3), H8/T10, m = 3")
            c = cf(zeros(1, 3), zeros(3, 3), 0, 0)
            return _gpu_bilform(assembler, self, geom, u, FORM_DOT, vec(Matrix{Float64}(c)'))
        end
    end
end

function bilform_dot(self::FEMM, assembler::SysmatAssemblerSparseHRZLumpingSymmHIP, geom::NodalField{FT},
                     u::NodalField{T}, cf::DC; m = 3) where {FEMM<:AbstractFEMM,FT,T,DC<:DataCache}
    (_onpath(self, cf) && m == 3) || error("the lumped path needs a constant DataCache, CSys(3), H8/T10, m = 3")
    c = cf(zeros(1, 3), zeros(3, 3), 0, 0)
    coeff = vec(Matrix{Float64}(c)')
    kind, order = _rule(self.integdomain.integration_rule)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:feasm_bilform_dot_lumped, libfeasm), Cint,
                (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}, Cint, Ptr{Ptr{Cvoid}}),
                _mesh_for(self, geom, u), kind, order, coeff, length(coeff), h), context())
    return _tocsc(h[])
end

"""
    SysvecAssemblerHIP()

In place of `SysvecAssembler` (src/AssemblyModule.jl:853-917).  The protocol
(startassembly!/assemble!/makevector!) keeps the reference's host buffer;
`linform_dot` with a constant force on H8/T10 solids (m = 3) runs on the GPU.
"""
mutable struct SysvecAssemblerHIP <: AbstractSysvecAssembler
    _buffer::Vector{Float64}
    _row_nalldofs::Int
end
SysvecAssemblerHIP(z::Float64 = 0.0) = SysvecAssemblerHIP(Float64[z], 1)

function startassembly!(self::SysvecAssemblerHIP, row_nalldofs::IT) where {IT<:Integer}
    self._row_nalldofs = row_nalldofs
    resize!(self._buffer, row_nalldofs)
    fill!(self._buffer, 0.0)
    return self
end

function assemble!(self::SysvecAssemblerHIP, vec::MV, dofnums::IV) where {MV,IV}
    for i in eachindex(dofnums)
        gi = dofnums[i]
        gi < 1 && error("Row degree of freedom < 1")
        gi > self._row_nalldofs && error("Row degree of freedom > size")
        self._buffer[gi] += vec[i]
    end
end

makevector!(self::SysvecAssemblerHIP) = deepcopy(self._buffer)

function linform_dot(self::FEMM, assembler::SysvecAssemblerHIP, geom::NodalField{FT}, P::NodalField{T},
                     f::DC, m) where {FEMM<:AbstractFEMM,FT,T,DC<:DataCache}
    if !(_onpath(self, f) && m == 3)
        return invoke(linform_dot, Tuple{AbstractFEMM,AbstractSysvecAssembler,NodalField,NodalField,DataCache,Any},
                      self, assembler, geom, P, f, m)
    end
    force = Vector{Float64}(vec(f(zeros(1, 3), zeros(3, 3), 0, 0)))
    kind, order = _rule(self.integdomain.integration_rule)
    startassembly!(assembler, nalldofs(P))
    check(ccall((:feasm_linform_dot, libfeasm), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}, Cint, Ptr{Float64}),
                _mesh_for(self, geom, P), kind, order, force, length(force), assembler._buffer), context())
    return makevector!(assembler)
end

export SysmatAssemblerSparseHIP, SysmatAssemblerSparseSymmHIP, SysmatAssemblerSparseHRZLumpingSymmHIP,
       SysmatAssemblerFFBlockHIP, SysvecAssemblerHIP

end # module
