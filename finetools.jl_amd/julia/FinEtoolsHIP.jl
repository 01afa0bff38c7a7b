"""
    FinEtoolsHIP

Julia side of the drop-in: assemblers compatible with FinEtools'
`SysmatAssemblerSparse`, `SysmatAssemblerSparseSymm`, `SysmatAssemblerSparseDiag`,
`SysmatAssemblerSparseHRZLumpingSymm`, `SysmatAssemblerFFBlock` and
`SysvecAssembler`, backed by libfeasm.so, plus `bilform_*` / `linform_dot`
methods specialised on them so that the whole element loop of
`bilform_diffusion`, `bilform_lin_elastic`, `bilform_dot` and `linform_dot`
runs on an MI355X, with the reference's call signatures.

    using FinEtools, FinEtoolsHIP
    a = SysmatAssemblerSparseHIP(0.0)
    K = bilform_lin_elastic(femm, a, geom, u, DeforModelRed3D, DataCache(D))
    K = bilform_lin_elastic(femm, geom, u, DeforModelRed3D, DataCache(D))   # default assembler, also on the GPU

Generic protocol calls (startassembly!/assemble!/makematrix!) on the same
assemblers buffer the triples as the reference's methods do and build the CSC
with the device sparse() (for the symmetric assembler then S + transpose(S)
on the device).  Anything the device element loop does not cover
(non-constant DataCache, a material frame other than the identity, other
element types/rules) falls back to FinEtools' own generic method through
`invoke`.

Device meshes are cached per assembler, keyed by the identity of the element
set, the geometry field and the unknown field, and released by the
assembler's finalizer.  Moved nodes are picked up when
`coordinates_changed!(geom)` is called (or when a sampled fingerprint of the
coordinates changes); a renumbered field when `dofs_changed!(u)` is called
(or its sampled fingerprint changes).

This file is not exercised in the build container (no Julia there); the C ABI
it binds is tested from Python (tests/).
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
const ASM_SPARSE, ASM_SYMM, ASM_DIAG, ASM_HRZ = Cint(0), Cint(1), Cint(2), Cint(3)

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

_destroy_matrix(h::Ptr{Cvoid}) = (h == C_NULL || ccall((:feasm_matrix_destroy, libfeasm), Cint, (Ptr{Cvoid},), h); nothing)
_destroy_mesh(h::Ptr{Cvoid}) = (h == C_NULL || ccall((:feasm_mesh_destroy, libfeasm), Cint, (Ptr{Cvoid},), h); nothing)

# ---------------------------------------------------------------- change tracking
# Version counters bumped by the user after changing a field in place; a cached
# device mesh records the versions it was built from.
const _VERSIONS = WeakKeyDict{Any,Int}()
_version(x) = get(_VERSIONS, x, 0)

"""
    coordinates_changed!(geom)

Tell the device caches that `geom.values` changed in place (moved nodes): the
next GPU assembly uploads the new coordinates and keeps the pattern.
"""
coordinates_changed!(geom) = (_VERSIONS[geom] = _version(geom) + 1; geom)

"""
    dofs_changed!(u)

Tell the device caches that `u.dofnums` changed in place (a new numbering).
"""
dofs_changed!(u) = (_VERSIONS[u] = _version(u) + 1; u)

# O(1) fingerprint: sizes, the data pointer and a strided sample of at most
# 256 entries (catches most in-place changes without reading whole arrays)
function _fingerprint(A::AbstractArray)
    n = length(A)
    h = hash((size(A), UInt(pointer(A))))
    n == 0 && return h
    step = max(1, n ÷ 256)
    @inbounds for i in 1:step:n
        h = hash(A[i], h)
    end
    return hash(A[n], h)
end

# ---------------------------------------------------------------- device mesh cache
mutable struct CachedMesh
    mesh::Ptr{Cvoid}
    conn_stamp::UInt   # fingerprint of the connectivity (object ids can be re-used after GC)
    geom_stamp::Tuple{Int,UInt}
    dofs_stamp::Tuple{Int,UInt}
    mats::Dict{Int,Ptr{Cvoid}}   # re-used device matrices by DOFs per node (feasm_bilform_into)
end

const MeshKey = Tuple{UInt,UInt,UInt}

function _release!(cache::Dict{MeshKey,CachedMesh})
    for c in values(cache)
        foreach(_destroy_matrix, values(c.mats))
        _destroy_mesh(c.mesh)
    end
    empty!(cache)
end

function _set_dofs(mesh::Ptr{Cvoid}, u)
    dn = Matrix{Int64}(u.dofnums)
    check(ccall((:feasm_dofs_set, libfeasm), Cint, (Ptr{Cvoid}, Cint, Ptr{Int64}, Int64),
                mesh, size(dn, 2), dn, nalldofs(u)), context())
end

function _mesh_for!(cache::Dict{MeshKey,CachedMesh}, femm, geom, u)
    fes = finite_elements(femm)
    key = (objectid(fes), objectid(geom), objectid(u))
    gs = (_version(geom), _fingerprint(geom.values))
    ds = (_version(u), _fingerprint(u.dofnums))
    cs = _fingerprint(fes.conn)
    c = get(cache, key, nothing)
    if c !== nothing && c.conn_stamp != cs   # another element set under a re-used id
        foreach(_destroy_matrix, values(c.mats))
        _destroy_mesh(c.mesh)
        delete!(cache, key)
        c = nothing
    end
    if c !== nothing
        if c.geom_stamp != gs   # moved nodes: new coordinates, same pattern and tiles
            xyz = Matrix{Float64}(permutedims(geom.values))   # row-major nnodes x 3
            check(ccall((:feasm_mesh_set_xyz, libfeasm), Cint, (Ptr{Cvoid}, Ptr{Float64}), c.mesh, xyz), context())
            c.geom_stamp = gs
        end
        if c.dofs_stamp != ds   # a new numbering: new pattern, the cached matrices go
            foreach(_destroy_matrix, values(c.mats))
            empty!(c.mats)
            _set_dofs(c.mesh, u)
            c.dofs_stamp = ds
        end
        return c
    end
    conn = Int64[c for e in fes.conn for c in e]                 # nelem x npe, row-major
    xyz = Matrix{Float64}(geom.values)                            # nnodes x 3, column-major
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:feasm_mesh_create, libfeasm), Cint,
                (Ptr{Cvoid}, Cint, Int64, Cint, Ptr{Int64}, Int64, Ptr{Float64}, Ptr{Ptr{Cvoid}}),
                context(), _etype(fes), count(fes), nodesperelem(fes), conn, size(xyz, 1), xyz, h), context())
    _set_dofs(h[], u)
    c = CachedMesh(h[], cs, gs, ds, Dict{Int,Ptr{Cvoid}}())
    cache[key] = c
    return c
end

# ---------------------------------------------------------------- assemblers
abstract type HIPAssembler <: AbstractSysmatAssembler end

# The buffered-triples assemblers share one layout: the kind selects the
# reference methods the C library follows.
for (T, kind, ref) in ((:SysmatAssemblerSparseHIP, ASM_SPARSE, "SysmatAssemblerSparse (src/AssemblyModule.jl:88-329)"),
                       (:SysmatAssemblerSparseSymmHIP, ASM_SYMM, "SysmatAssemblerSparseSymm (src/AssemblyModule.jl:342-583)"),
                       (:SysmatAssemblerSparseDiagHIP, ASM_DIAG, "SysmatAssemblerSparseDiag (src/AssemblyModule.jl:599-794)"),
                       (:SysmatAssemblerSparseHRZLumpingSymmHIP, ASM_HRZ,
                        "SysmatAssemblerSparseHRZLumpingSymm (src/AssemblyModule.jl:943-1141)"))
    doc = "    $(T)(z::Float64 = 0.0, nomatrixresult = false)\n\nIn place of $(ref): the generic " *
          "protocol buffers the triples as the reference does and makematrix! builds the CSC on the device; " *
          "the bilform_* methods specialised on it run the element loop on the device."
    @eval begin
        @doc $doc
        mutable struct $T <: HIPAssembler
            _matbuffer::Vector{Float64}   # kept so that eltype(a) works (AssemblyModule.jl:27)
            _coo::Ptr{Cvoid}
            _nomatrixresult::Bool
            _force_init::Bool
            _row_nalldofs::Int
            _col_nalldofs::Int
            _meshes::Dict{MeshKey,CachedMesh}
        end
        function $T(z::Float64 = 0.0, nomatrixresult = false)
            h = Ref{Ptr{Cvoid}}(C_NULL)
            check(ccall((:feasm_coo_create_kind, libfeasm), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{Ptr{Cvoid}}),
                        context(), $kind, nomatrixresult, h), context())
            a = $T(Float64[z], h[], nomatrixresult, false, 0, 0, Dict{MeshKey,CachedMesh}())
            finalizer(a) do x
                ccall((:feasm_coo_destroy, libfeasm), Cint, (Ptr{Cvoid},), x._coo)
                _release!(x._meshes)
            end
            return a
        end
    end
end

"""
    SysmatAssemblerFFBlockHIP(row_nfreedofs, col_nfreedofs = row_nfreedofs)

In place of `SysmatAssemblerFFBlock` (src/AssemblyModule.jl:1149-1231): the
protocol delegates to a wrapped `SysmatAssemblerSparseHIP`, and makematrix!
extracts the leading free-free block (`matrix_blocked_ff`,
src/MatrixUtilityModule.jl:675-690) on the device.
"""
mutable struct SysmatAssemblerFFBlockHIP <: AbstractSysmatAssembler
    _a::SysmatAssemblerSparseHIP
    _row_nfreedofs::Int
    _col_nfreedofs::Int
end
SysmatAssemblerFFBlockHIP(row_nfreedofs::Integer, col_nfreedofs::Integer = row_nfreedofs) =
    SysmatAssemblerFFBlockHIP(SysmatAssemblerSparseHIP(0.0), row_nfreedofs, col_nfreedofs)
Base.eltype(self::SysmatAssemblerFFBlockHIP) = eltype(self._a)

# ---------------------------------------------------------------- generic protocol
function startassembly!(self::HIPAssembler, elem_mat_nrows::IT, elem_mat_ncols::IT, n_elem_mats::IT,
                        row_nalldofs::IT, col_nalldofs::IT; force_init = false) where {IT<:Integer}
    check(ccall((:feasm_coo_start, libfeasm), Cint, (Ptr{Cvoid}, Int64, Int64, Int64, Int64, Int64),
                self._coo, elem_mat_nrows, elem_mat_ncols, n_elem_mats, row_nalldofs, col_nalldofs), context())
    if self._row_nalldofs == 0
        self._row_nalldofs, self._col_nalldofs = row_nalldofs, col_nalldofs
    end
    self._force_init = force_init
    return self
end

function assemble!(self::HIPAssembler, mat::AbstractMatrix, dofnums_row, dofnums_col)
    m = Matrix{Float64}(mat)
    r = Vector{Int64}(vec(dofnums_row))
    c = Vector{Int64}(vec(dofnums_col))
    check(ccall((:feasm_coo_assemble, libfeasm), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Int64, Int64, Ptr{Int64}, Int64, Ptr{Int64}, Int64),
                self._coo, m, size(m, 1), size(m, 2), r, length(r), c, length(c)), context())
    return self
end

# copy a device matrix out as Julia's SparseMatrixCSC{Float64,Int}
function _copy_csc(h::Ptr{Cvoid})
    m, n, nnz = Ref{Int64}(0), Ref{Int64}(0), Ref{Int64}(0)
    check(ccall((:feasm_matrix_dims, libfeasm), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}, Ref{Int64}),
                h, m, n, nnz), context())
    colptr = Vector{Int64}(undef, n[] + 1)
    rowval = Vector{Int64}(undef, nnz[])
    nzval = Vector{Float64}(undef, nnz[])
    check(ccall((:feasm_matrix_copy_csc, libfeasm), Cint, (Ptr{Cvoid}, Cint, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}),
                h, 1, colptr, rowval, nzval), context())
    return SparseMatrixCSC{Float64,Int64}(m[], n[], colptr, rowval, nzval)
end

# copy out, then release the device matrix
_tocsc(h::Ptr{Cvoid}) = (S = _copy_csc(h); _destroy_matrix(h); S)

function _makematrix_handle(self::HIPAssembler)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:feasm_coo_set_nomatrixresult, libfeasm), Cint, (Ptr{Cvoid}, Cint), self._coo,
                self._nomatrixresult), context())
    check(ccall((:feasm_coo_makematrix, libfeasm), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}), self._coo, h), context())
    return h[]
end

function makematrix!(self::HIPAssembler)
    h = _makematrix_handle(self)
    h == C_NULL && return spzeros(self._row_nalldofs, self._col_nalldofs)
    return _tocsc(h)
end

function _ff_block(h::Ptr{Cvoid}, r, c)
    h2 = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:feasm_matrix_ff_block, libfeasm), Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{Ptr{Cvoid}}), h, r, c, h2)
    rc == 0 || _destroy_matrix(h)
    check(rc, context())
    return h2[]
end

startassembly!(self::SysmatAssemblerFFBlockHIP, elem_mat_nrows::IT, elem_mat_ncols::IT, n_elem_mats::IT,
               row_nalldofs::IT, col_nalldofs::IT; force_init = false) where {IT<:Integer} =
    (startassembly!(self._a, elem_mat_nrows, elem_mat_ncols, n_elem_mats, row_nalldofs, col_nalldofs;
                    force_init = force_init); self)
assemble!(self::SysmatAssemblerFFBlockHIP, mat::AbstractMatrix, dofnums_row, dofnums_col) =
    (assemble!(self._a, mat, dofnums_row, dofnums_col); self)

function makematrix!(self::SysmatAssemblerFFBlockHIP)
    h = _makematrix_handle(self._a)
    if h == C_NULL   # nomatrixresult: the block of the reference's spzeros
        return FinEtools.MatrixUtilityModule.matrix_blocked_ff(spzeros(self._a._row_nalldofs, self._a._col_nalldofs),
                                                               self._row_nfreedofs, self._col_nfreedofs)
    end
    h2 = _ff_block(h, self._row_nfreedofs, self._col_nfreedofs)
    _destroy_matrix(h)
    return _tocsc(h2)
end

# ---------------------------------------------------------------- GPU bilinear forms
_etype(::FESetH8) = FEASM_H8
_etype(::FESetT10) = FEASM_T10
_etype(::Any) = nothing
_rule(r::GaussRule) = (RULE_GAUSS, r.dim)   # GaussRule stores order in `dim` (IntegRuleModule.jl:391-392)
_rule(r::TetRule) = (RULE_TET, r.npts)
_rule(::Any) = nothing

# The element loop on the device into the cached matrix of this mesh and DOF
# count (feasm_bilform the first time, feasm_bilform_into afterwards: the
# pattern and the device values are re-used).  Returns the device handle,
# owned by the cache.
function _gpu_bilform_cached(cache, femm, geom, u, form, coeff::Vector{Float64})
    kind, order = _rule(femm.integdomain.integration_rule)
    c = _mesh_for!(cache, femm, geom, u)
    ndn = size(u.dofnums, 2)
    h = get(c.mats, ndn, C_NULL)
    if h != C_NULL
        check(ccall((:feasm_bilform_into, libfeasm), Cint,
                    (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Float64}, Cint, Cint, Ptr{Cvoid}),
                    c.mesh, form, kind, order, coeff, length(coeff), 0, h), context())
        return (h, true)
    end
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:feasm_bilform, libfeasm), Cint,
                (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Float64}, Cint, Cint, Ptr{Ptr{Cvoid}}),
                c.mesh, form, kind, order, coeff, length(coeff), 0, r), context())
    # a matrix in the mesh's resident pattern (node-major numbering) is kept
    # for re-assembly; a general numbering's matrix (COO path) is not
    resident = ccall((:feasm_pattern_info, libfeasm), Cint, (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}),
                     c.mesh, C_NULL, C_NULL) == 0
    resident && (c.mats[ndn] = r[])
    return (r[], resident)
end

# the assembled matrix as the assembler's makematrix! would return it
function _finish(::Union{SysmatAssemblerSparseHIP}, h, cached)
    return cached ? _copy_csc(h) : _tocsc(h)
end

function _finish(::SysmatAssemblerSparseSymmHIP, h, cached)
    h2 = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:feasm_matrix_dropzeros, libfeasm), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}), h, h2), context())
    cached || _destroy_matrix(h)
    return _tocsc(h2[])
end

function _finish(a::SysmatAssemblerFFBlockHIP, h, cached)
    h2 = _ff_block(h, a._row_nfreedofs, a._col_nfreedofs)
    cached || _destroy_matrix(h)
    return _tocsc(h2)
end

_cache(a::HIPAssembler) = a._meshes
_cache(a::SysmatAssemblerFFBlockHIP) = a._a._meshes

function _gpu_bilform(a, femm, geom, u, form, coeff::Vector{Float64})
    h, cached = _gpu_bilform_cached(_cache(a), femm, geom, u, form, coeff)
    return _finish(a, h, cached)
end

# The device path applies to a constant DataCache (its fill function is the
# constant closure of DataCache(data), DataCacheModule.jl:76-89), the identity
# material frame (CSys(3), CSysModule.jl:154-181) and H8/T10 with a
# Gauss/Tet rule; everything else defers to FinEtools' generic methods.
_constant(cf::DataCache) = occursin("fillcache_constant", string(nameof(typeof(cf._fillcache!))))
_onpath(femm, cf) = (_etype(finite_elements(femm)) !== nothing) &&
                    (_rule(femm.integdomain.integration_rule) !== nothing) &&
                    femm.mcsys.isidentity && _constant(cf)

const _ELEMENT_LOOP = Union{SysmatAssemblerSparseHIP,SysmatAssemblerSparseSymmHIP,SysmatAssemblerFFBlockHIP}

function bilform_diffusion(self::FEMM, assembler::_ELEMENT_LOOP, geom::NodalField{FT},
                           u::NodalField{T}, cf::DC) where {FEMM<:AbstractFEMM,FT,T,DC<:DataCache}
    if !_onpath(self, cf)
        return invoke(bilform_diffusion, Tuple{AbstractFEMM,AbstractSysmatAssembler,NodalField,NodalField,DataCache},
                      self, assembler, geom, u, cf)
    end
    c = cf(zeros(1, 3), zeros(3, 3), 0, 0)
    isempty(size(c)) && return _gpu_bilform(assembler, self, geom, u, FORM_DIFF_ISO, [Float64(c)])
    return _gpu_bilform(assembler, self, geom, u, FORM_DIFF_MAT, vec(Matrix{Float64}(c)'))
end

function bilform_lin_elastic(self::FEMM, assembler::_ELEMENT_LOOP, geom::NodalField{FT},
                             u::NodalField{T}, mr::Type{DeforModelRed3D}, cf::DC) where {FEMM<:AbstractFEMM,FT,T,DC<:DataCache}
    if !_onpath(self, cf)
        return invoke(bilform_lin_elastic,
                      Tuple{AbstractFEMM,AbstractSysmatAssembler,NodalField,NodalField,Type,DataCache},
                      self, assembler, geom, u, mr, cf)
    end
    D = cf(zeros(1, 3), zeros(3, 3), 0, 0)
    return _gpu_bilform(assembler, self, geom, u, FORM_LIN_ELASTIC, vec(Matrix{Float64}(D)'))
end

function bilform_dot(self::FEMM, assembler::_ELEMENT_LOOP, geom::NodalField{FT}, u::NodalField{T},
                     cf::DC; m = 3) where {FEMM<:AbstractFEMM,FT,T,DC<:DataCache}
    if !_onpath(self, cf) || m != 3
        return invoke(bilform_dot, Tuple{AbstractFEMM,AbstractSysmatAssembler,NodalField,NodalField,DataCache},
                      self, assembler, geom, u, cf; m = m)
    end
    c = cf(zeros(1, 3), zeros(3, 3), 0, 0)
    return _gpu_bilform(assembler, self, geom, u, FORM_DOT, vec(Matrix{Float64}(c)'))
end

function bilform_dot(self::FEMM, assembler::SysmatAssemblerSparseHRZLumpingSymmHIP, geom::NodalField{FT},
                     u::NodalField{T}, cf::DC; m = 3) where {FEMM<:AbstractFEMM,FT,T,DC<:DataCache}
    if !(_onpath(self, cf) && m == 3)
        return invoke(bilform_dot, Tuple{AbstractFEMM,AbstractSysmatAssembler,NodalField,NodalField,DataCache},
                      self, assembler, geom, u, cf; m = m)
    end
    c = cf(zeros(1, 3), zeros(3, 3), 0, 0)
    coeff = vec(Matrix{Float64}(c)')
    kind, order = _rule(self.integdomain.integration_rule)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:feasm_bilform_dot_lumped, libfeasm), Cint,
                (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}, Cint, Ptr{Ptr{Cvoid}}),
                _mesh_for!(assembler._meshes, self, geom, u).mesh, kind, order, coeff, length(coeff), h), context())
    return _tocsc(h[])
end

# ---------------------------------------------------------------- the default assembler
# FinEtools' convenience methods bilform_diffusion(femm, geom, u, cf),
# bilform_lin_elastic(femm, geom, u, mr, cf) and bilform_dot(femm, geom, u, cf)
# construct a SysmatAssemblerSparseSymm and call the five-argument method
# (src/FEMMBaseModule.jl:1368-1376, 1537-1545, 1815-1824).  The methods below
# are more specific than FinEtools' generic ones for that assembler type, so
# those calls reach the device too (through one module-wide symmetric HIP
# assembler, whose mesh cache they share).  `route_default_assembler(false)`
# turns this off; an assembler in nomatrixresult mode, or a call off the device
# path, goes to FinEtools' own method.
const ROUTE_DEFAULT = Ref(true)
route_default_assembler(flag::Bool) = (ROUTE_DEFAULT[] = flag)
const _DEFAULT_SYMM = Ref{Any}(nothing)
function _default_symm()
    _DEFAULT_SYMM[] === nothing && (_DEFAULT_SYMM[] = SysmatAssemblerSparseSymmHIP(0.0))
    return _DEFAULT_SYMM[]
end
_route(a::SysmatAssemblerSparseSymm, femm, cf) = ROUTE_DEFAULT[] && !a._nomatrixresult && _onpath(femm, cf)

function bilform_diffusion(self::FEMM, assembler::SysmatAssemblerSparseSymm, geom::NodalField{FT},
                           u::NodalField{T}, cf::DC) where {FEMM<:AbstractFEMM,FT,T,DC<:DataCache}
    _route(assembler, self, cf) && return bilform_diffusion(self, _default_symm(), geom, u, cf)
    return invoke(bilform_diffusion, Tuple{AbstractFEMM,AbstractSysmatAssembler,NodalField,NodalField,DataCache},
                  self, assembler, geom, u, cf)
end

function bilform_lin_elastic(self::FEMM, assembler::SysmatAssemblerSparseSymm, geom::NodalField{FT},
                             u::NodalField{T}, mr::Type{DeforModelRed3D}, cf::DC) where {FEMM<:AbstractFEMM,FT,T,DC<:DataCache}
    _route(assembler, self, cf) && return bilform_lin_elastic(self, _default_symm(), geom, u, mr, cf)
    return invoke(bilform_lin_elastic, Tuple{AbstractFEMM,AbstractSysmatAssembler,NodalField,NodalField,Type,DataCache},
                  self, assembler, geom, u, mr, cf)
end

function bilform_dot(self::FEMM, assembler::SysmatAssemblerSparseSymm, geom::NodalField{FT}, u::NodalField{T},
                     cf::DC; m = 3) where {FEMM<:AbstractFEMM,FT,T,DC<:DataCache}
    (_route(assembler, self, cf) && m == 3) && return bilform_dot(self, _default_symm(), geom, u, cf; m = m)
    return invoke(bilform_dot, Tuple{AbstractFEMM,AbstractSysmatAssembler,NodalField,NodalField,DataCache},
                  self, assembler, geom, u, cf; m = m)
end

# ---------------------------------------------------------------- load vectors
"""
    SysvecAssemblerHIP()

In place of `SysvecAssembler` (src/AssemblyModule.jl:853-917).  The protocol
(startassembly!/assemble!/makevector!) keeps the reference's host buffer;
`linform_dot` with a constant force on H8/T10 solids (m = 3) runs on the GPU.
"""
mutable struct SysvecAssemblerHIP <: AbstractSysvecAssembler
    _buffer::Vector{Float64}
    _row_nalldofs::Int
    _meshes::Dict{MeshKey,CachedMesh}
end
function SysvecAssemblerHIP(z::Float64 = 0.0)
    a = SysvecAssemblerHIP(Float64[z], 1, Dict{MeshKey,CachedMesh}())
    finalizer(x -> _release!(x._meshes), a)
    return a
end

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
                _mesh_for!(assembler._meshes, self, geom, P).mesh, kind, order, force, length(force),
                assembler._buffer), context())
    return makevector!(assembler)
end

export SysmatAssemblerSparseHIP, SysmatAssemblerSparseSymmHIP, SysmatAssemblerSparseDiagHIP,
       SysmatAssemblerSparseHRZLumpingSymmHIP, SysmatAssemblerFFBlockHIP, SysvecAssemblerHIP,
       coordinates_changed!, dofs_changed!, route_default_assembler

end # module
