# Julia-side binding of libsds_hip: a package extension that routes the numeric and
# solve phases of SparseDirectSolver.jl to the MI355X library through `ccall`.
#
# It hooks the three functions the reference dispatches its numeric phase to
# (src/solver.jl:717-741) and the solve phase (src/solver.jl:810-871) for solvers
# whose values live in AMDGPU `ROCArray`s. The analysis (`Symbolic`) stays the
# reference's own; its device maps (`DEVICE_MAPS`, src/symbolic/maps.jl:16-18) are
# passed once to `sds_upload_symbolic`.
#
# This file is not loaded by the Python tests (there is no Julia toolchain in the build
# image); it is the binding a maintainer adds to the reference's `ext/` directory.
module SparseDirectSolverMI355XExt

using SparseDirectSolver
using SparseDirectSolver: DirectSolver, Symbolic, Phase, PHASE_SOLVE, InvalidValueError,
                          NotSupportedError, FactorizationError
using AMDGPU

const libsds_hip = get(ENV, "SDS_HIP_LIB", "libsds_hip")

# status codes of include/sds_hip.h
const SDS_SUCCESS = Cint(0)
const SDS_INVALID_VALUE = Cint(1)
const SDS_NOT_SUPPORTED = Cint(2)
const SDS_FACTORIZATION_ERROR = Cint(3)
const SDS_R64 = Cint(0)
const SDS_R32 = Cint(1)

# sds_symbolic_desc (include/sds_hip.h), field for field
struct SymbolicDesc
    n::Int64
    nnz::Int64
    ns::Int64
    structure::Int32
    index_bits::Int32
    perm::Ptr{Cvoid}
    super_ptr::Ptr{Cvoid}
    snparent::Ptr{Cvoid}
    rowptr::Ptr{Cvoid}
    rowval::Ptr{Cvoid}
    front_ptr::Ptr{Cvoid}
    front_nrows::Ptr{Cvoid}
    front_ncols::Ptr{Cvoid}
    child_ptr::Ptr{Cvoid}
    child_list::Ptr{Cvoid}
    relind_ptr::Ptr{Cvoid}
    relind::Ptr{Cvoid}
    amap::Ptr{Cvoid}
    amap_ptr::Ptr{Cvoid}
    amap_src::Ptr{Cvoid}
    full_nrows::Int64
    full_nnz::Int64
    full_rowptr::Ptr{Cvoid}
    full_colval::Ptr{Cvoid}
    full_src::Ptr{Cvoid}
end

# sds_factor_params (include/sds_hip.h)
struct FactorParams
    pivot_type::Int32
    pivot_epsilon_alg::Int32
    pivot_threshold::Float64
    pivot_epsilon::Float64
    pivot_sign::Ptr{Int8}
    interrupt::Ptr{Int32}
end

mutable struct Handle
    ptr::Ptr{Cvoid}
end

function _check(h::Handle, st::Cint)
    st == SDS_SUCCESS && return nothing
    msg = unsafe_string(ccall((:sds_last_error, libsds_hip), Cstring, (Ptr{Cvoid},), h.ptr))
    st == SDS_INVALID_VALUE && throw(InvalidValueError(msg))
    st == SDS_NOT_SUPPORTED && throw(NotSupportedError(msg))
    throw(FactorizationError(msg))
end

function create_handle(stream::AMDGPU.HIPStream)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    st = ccall((:sds_create, libsds_hip), Cint, (Ptr{Ptr{Cvoid}}, Ptr{Cvoid}), r, stream.stream)
    st == SDS_SUCCESS || throw(InvalidValueError("sds_create failed ($st)"))
    h = Handle(r[])
    finalizer(x -> ccall((:sds_destroy, libsds_hip), Cint, (Ptr{Cvoid},), x.ptr), h)
    return h
end

# The host copies of the device maps (1-based, as the Symbolic holds them); the
# library converts them once (replaces `adapt(backend, S, INT)`, src/symbolic/maps.jl:288-294).
function upload_symbolic!(h::Handle, S::Symbolic{INT}, full) where {INT}
    host(v) = Array(v)
    keep = map(host, (S.perm, S.super_ptr, S.snparent, S.rowptr, S.rowval, S.front_ptr, S.front_nrows,
                      S.front_ncols, S.child_ptr, S.child_list, S.relind_ptr, S.relind, S.amap, S.amap_ptr,
                      S.amap_src))
    frp, fcv, fsrc = full   # the refinement map of src/solve/refinement.jl:67-101
    GC.@preserve keep frp fcv fsrc begin
        p = map(pointer, keep)
        d = SymbolicDesc(S.n, S.nnz, length(S.super_ptr) - 1, Int32(S.structure), Int32(8 * sizeof(INT)),
                         p..., Int64(length(frp) - 1), Int64(length(fcv)), pointer(frp), pointer(fcv),
                         pointer(fsrc))
        _check(h, ccall((:sds_upload_symbolic, libsds_hip), Cint, (Ptr{Cvoid}, Ref{SymbolicDesc}), h.ptr, d))
    end
    return nothing
end

alloc_numeric!(h::Handle, ::Type{T}) where {T} =
    _check(h, ccall((:sds_alloc_numeric, libsds_hip), Cint, (Ptr{Cvoid}, Cint, Cint), h.ptr,
                    T === Float32 ? SDS_R32 : SDS_R64, Cint(1)))

# replaces factorize_cholesky! / factorize_ldlt! / factorize_lu! (src/numeric/factorize.jl:315-360)
function factorize!(h::Handle, nzval::ROCVector{Float64}, opts)
    info = Ref{Int64}(0)
    psign = opts.pivot_sign === nothing ? C_NULL : pointer(opts.pivot_sign)
    p = FactorParams(Int32(opts.pivot_type), Int32(opts.pivot_epsilon_alg), opts.pivot_threshold,
                     opts.pivot_epsilon, psign, C_NULL)
    _check(h, ccall((:sds_factorize, libsds_hip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ref{FactorParams}, Ref{Int64}),
                    h.ptr, pointer(nzval), p, info))
    return Int(info[])
end

# replaces the body of _solve_phase! (src/solver.jl:846-869): phase bits of src/types.jl:56-71
function solve!(h::Handle, phase::Phase, X::ROCMatrix{Float64}, B::ROCMatrix{Float64},
                nzval::ROCVector{Float64}, opts)
    steps = Ref{Int32}(0)
    # ir_mode "ir" / "fgmres" (src/types.jl:306): IR_PLAIN = 0, IR_FGMRES = 1 in the C ABI
    _check(h, ccall((:sds_set_ir_mode, libsds_hip), Cint, (Ptr{Cvoid}, Cint), h.ptr,
                    Cint(opts.ir_mode == IR_FGMRES ? 1 : 0)))
    _check(h, ccall((:sds_solve, libsds_hip), Cint,
                    (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Cint, Ptr{Float64}, Cint,
                     Float64, Ref{Int32}),
                    h.ptr, Cint(phase), Cint(opts.solve_mode), pointer(B), stride(B, 2), pointer(X), stride(X, 2),
                    Cint(size(B, 2)), pointer(nzval), Cint(opts.ir_n_steps), opts.ir_tol, steps))
    return Int(steps[])
end

end # module
