# Julia-side binding of libsds_hip: a package extension that routes the numeric and
# solve phases of SparseDirectSolver.jl to the MI355X library through `ccall`.
#
# It provides the calls that replace the three functions the reference dispatches its
# numeric phase to (src/solver.jl:717-741) and the solve phase (src/solver.jl:810-871) for
# solvers whose values live in AMDGPU `ROCArray`s, plus the library's own additions (HIP
# graph replay, tile autotuning, the adjoint scatter and a ChainRulesCore `rrule` of the
# solve). The analysis (`Symbolic`) stays the reference's own; its device maps
# (`DEVICE_MAPS`, src/symbolic/maps.jl:16-18) are passed once to `sds_upload_symbolic`.
#
# This file is not loaded by the Python tests (there is no Julia toolchain in the build
# image); it is the binding a maintainer adds to the reference's `ext/` directory, with
# AMDGPU and ChainRulesCore as weak dependencies. tests/test_julia_binding.py checks every
# `ccall` below against include/sds_hip.h (symbol, argument count, return type).
module SparseDirectSolverMI355XExt

using SparseDirectSolver
using SparseDirectSolver: DirectSolver, Symbolic, Phase, PHASE_SOLVE, IR_FGMRES, InvalidValueError,
                          NotSupportedError, FactorizationError, InterruptedError
using AMDGPU
using ChainRulesCore

const libsds_hip = get(ENV, "SDS_HIP_LIB", "libsds_hip")

# status codes of include/sds_hip.h
const SDS_SUCCESS = Cint(0)
const SDS_INVALID_VALUE = Cint(1)
const SDS_NOT_SUPPORTED = Cint(2)
const SDS_FACTORIZATION_ERROR = Cint(3)
const SDS_DEVICE_ERROR = Cint(4)
const SDS_INTERRUPTED = Cint(5)
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

# sds_factor_params (include/sds_hip.h); `interrupt` points at the Bool inside the
# reference's `user_host_interrupt::Threads.Atomic{Bool}` (src/options.jl:201)
struct FactorParams
    pivot_type::Int32
    pivot_epsilon_alg::Int32
    pivot_threshold::Float64
    pivot_epsilon::Float64
    pivot_sign::Ptr{Int8}
    interrupt::Ptr{UInt8}
end

mutable struct Handle
    ptr::Ptr{Cvoid}
end

function _check(h::Handle, st::Cint)
    st == SDS_SUCCESS && return nothing
    msg = unsafe_string(ccall((:sds_last_error, libsds_hip), Cstring, (Ptr{Cvoid},), h.ptr))
    st == SDS_INVALID_VALUE && throw(InvalidValueError(msg))
    st == SDS_NOT_SUPPORTED && throw(NotSupportedError(msg))
    st == SDS_INTERRUPTED && throw(InterruptedError(msg))
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

set_stream!(h::Handle, stream::AMDGPU.HIPStream) =
    _check(h, ccall((:sds_set_stream, libsds_hip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), h.ptr, stream.stream))

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

# Schur complement mode (src/solver.jl:400-416): the reference's analysis with schur_mode = 1
# orders the Schur indices last as the root supernode; the library then assembles that root
# without factoring it. Before alloc_numeric!.
set_schur!(h::Handle, nschur::Integer) =
    _check(h, ccall((:sds_set_schur, libsds_hip), Cint, (Ptr{Cvoid}, Int64), h.ptr, Int64(nschur)))

# Matching and scaling (matching_alg, src/solver.jl:543-604): the reference's host Matching `m`
# (compute_matching, src/matching/mc64.jl:370-392) and its entry weights (entry_scaling,
# :463-470); "G" passes the column permutation m.perm, the symmetric structures nothing. After
# upload_symbolic! of the matched analysis; the full-pattern map must be that of A itself.
function set_matching!(h::Handle, w::Vector{Float64}, m)
    rs, cs = Vector{Float64}(m.rscale), Vector{Float64}(m.cscale)
    q = m.symmetric ? nothing : Vector{Int64}(m.perm)
    GC.@preserve w rs cs q begin
        _check(h, ccall((:sds_set_matching, libsds_hip), Cint,
                        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Int64}),
                        h.ptr, pointer(w), pointer(rs), pointer(cs), q === nothing ? C_NULL : pointer(q)))
    end
    return nothing
end

# getparam(solver, "schur_matrix") into a dense device matrix (schur_matrix!, src/numeric/schur.jl:112-150)
function schur_matrix!(S::ROCMatrix{Float64}, h::Handle)
    _check(h, ccall((:sds_schur_export, libsds_hip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64), h.ptr, pointer(S),
                    stride(S, 2)))
    return S
end

alloc_numeric!(h::Handle, ::Type{T}) where {T} =
    _check(h, ccall((:sds_alloc_numeric, libsds_hip), Cint, (Ptr{Cvoid}, Cint, Cint), h.ptr,
                    T === Float32 ? SDS_R32 : SDS_R64, Cint(1)))

# The factor parameters of `opts` with raw pointers into `opts.pivot_sign` and the interrupt flag:
# call only inside `GC.@preserve opts`.
function _params(opts)
    psign = opts.pivot_sign === nothing ? Ptr{Int8}(C_NULL) : pointer(opts.pivot_sign)
    flag = opts.user_host_interrupt
    intr = flag === nothing ? Ptr{UInt8}(C_NULL) : Ptr{UInt8}(pointer_from_objref(flag))
    return FactorParams(Int32(opts.pivot_type), Int32(opts.pivot_epsilon_alg), opts.pivot_threshold,
                        opts.pivot_epsilon, psign, intr)
end

# replaces factorize_cholesky! / factorize_ldlt! / factorize_lu! (src/numeric/factorize.jl:315-360);
# the library polls the interrupt flag before every launch group (src/numeric/factorize.jl:365-379)
function factorize!(h::Handle, nzval::ROCVector{Float64}, opts)
    info = Ref{Int64}(0)
    GC.@preserve opts begin
        p = _params(opts)
        _check(h, ccall((:sds_factorize, libsds_hip), Cint,
                        (Ptr{Cvoid}, Ptr{Float64}, Ref{FactorParams}, Ref{Int64}), h.ptr, pointer(nzval), p, info))
    end
    return Int(info[])
end

set_ir_mode!(h::Handle, opts) =
    _check(h, ccall((:sds_set_ir_mode, libsds_hip), Cint, (Ptr{Cvoid}, Cint), h.ptr,
                    Cint(opts.ir_mode == IR_FGMRES ? 1 : 0)))

# replaces the body of _solve_phase! (src/solver.jl:846-869): phase bits of src/types.jl:56-71.
# `solve_mode` 0 solves A X = B, 1 A^T X = B (src/solve/refinement.jl:10-15).
function solve!(h::Handle, phase::Phase, X::ROCMatrix{Float64}, B::ROCMatrix{Float64},
                nzval::ROCVector{Float64}, opts; solve_mode::Integer = opts.solve_mode)
    steps = Ref{Int32}(0)
    set_ir_mode!(h, opts)
    _check(h, ccall((:sds_solve, libsds_hip), Cint,
                    (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Cint, Ptr{Float64}, Cint,
                     Float64, Ref{Int32}),
                    h.ptr, Cint(phase), Cint(solve_mode), pointer(B), stride(B, 2), pointer(X), stride(X, 2),
                    Cint(size(B, 2)), pointer(nzval), Cint(opts.ir_n_steps), opts.ir_tol, steps))
    return Int(steps[])
end

# the library's tile-shape autotuner: {extend-add rows, K chunk, trailing split, two-per-CU steps,
# small-front kernels, subtree KiB, panel block, pivot panels}; the numeric storage is reallocated,
# so factorize again afterwards
function autotune!(h::Handle, nzval::ROCVector{Float64}, opts; reps::Integer = 3)
    tiles = zeros(Int32, 8)
    GC.@preserve opts tiles begin
        p = _params(opts)
        _check(h, ccall((:sds_autotune, libsds_hip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ref{FactorParams}, Cint,
                                                           Ptr{Int32}),
                        h.ptr, pointer(nzval), p, Cint(reps), pointer(tiles)))
    end
    return tiles
end

# one HIP graph of refactorization + solve; every replay reads the current nzval and B, writes X
function graph_capture!(h::Handle, nzval::ROCVector{Float64}, X::ROCMatrix{Float64}, B::ROCMatrix{Float64},
                        opts; phase::Phase = PHASE_SOLVE, solve_mode::Integer = opts.solve_mode)
    set_ir_mode!(h, opts)
    GC.@preserve opts begin
        p = _params(opts)
        _check(h, ccall((:sds_graph_capture, libsds_hip), Cint,
                        (Ptr{Cvoid}, Ptr{Float64}, Ref{FactorParams}, Cint, Cint, Ptr{Float64}, Int64, Ptr{Float64},
                         Int64, Cint, Cint),
                        h.ptr, pointer(nzval), p, Cint(phase), Cint(solve_mode), pointer(B), stride(B, 2), pointer(X),
                        stride(X, 2), Cint(size(B, 2)), Cint(opts.ir_n_steps)))
    end
    return nothing
end

graph_launch!(h::Handle) = _check(h, ccall((:sds_graph_launch, libsds_hip), Cint, (Ptr{Cvoid},), h.ptr))

function graph_info(h::Handle)
    info = Ref{Int64}(0)
    _check(h, ccall((:sds_graph_info, libsds_hip), Cint, (Ptr{Cvoid}, Ref{Int64}), h.ptr, info))
    return Int(info[])
end

graph_destroy!(h::Handle) = _check(h, ccall((:sds_graph_destroy, libsds_hip), Cint, (Ptr{Cvoid},), h.ptr))

# grad[p] = -sum_r lambda[i, r] X[j, r] on the stored entries (i, j) of A (both positions of a
# mirrored symmetric entry); lambda = A^-T Xbar from a transposed solve on the same factor
function adjoint_scatter!(grad::ROCVector{Float64}, h::Handle, lambda::ROCMatrix{Float64}, X::ROCMatrix{Float64})
    _check(h, ccall((:sds_adjoint_scatter, libsds_hip), Cint,
                    (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Cint, Ptr{Float64}),
                    h.ptr, pointer(lambda), stride(lambda, 2), pointer(X), stride(X, 2), Cint(size(X, 2)),
                    pointer(grad)))
    return grad
end

# getparam data after a factorization (src/solver.jl getparam; the `Numeric` fields of
# src/numeric/storage.jl:161-179): per-front stats (npos, nneg, nzero, nperturbed, n2x2, info)
# and their totals; inertia = (totals[1], totals[2]), npivots = totals[4]
function stats(h::Handle, ns::Integer)
    st = zeros(Int64, 6 * ns)
    tot = zeros(Int64, 6)
    GC.@preserve st tot _check(h, ccall((:sds_get_stats, libsds_hip), Cint, (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}),
                                        h.ptr, pointer(st), pointer(tot)))
    return reshape(st, 6, ns), tot
end

# "perm_row" pieces: the local pivot order (1-based factor columns) and the pivot kinds
function pivots(h::Handle, n::Integer)
    piv = zeros(Int32, n)
    kind = zeros(Int8, n)
    GC.@preserve piv kind _check(h, ccall((:sds_get_pivots, libsds_hip), Cint, (Ptr{Cvoid}, Ptr{Int32}, Ptr{Int8}),
                                          h.ptr, pointer(piv), pointer(kind)))
    return piv, kind
end

# "diag": D's diagonal and its 2x2 subdiagonals (2n entries of the factor type)
function diag_blocks(h::Handle, n::Integer, ::Type{T}) where {T}
    d = zeros(T, 2n)
    GC.@preserve d _check(h, ccall((:sds_get_diag, libsds_hip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), h.ptr, pointer(d)))
    return d
end

# "memory_estimates" (src/symbolic/maps.jl:338-356), bytes
function memory_estimates(h::Handle)
    est = zeros(Int64, 16)
    GC.@preserve est _check(h, ccall((:sds_memory_estimates, libsds_hip), Cint, (Ptr{Cvoid}, Ptr{Int64}), h.ptr,
                                     pointer(est)))
    return est
end

# A factorized handle with the values it was factorized with: the differentiable solve.
struct FactoredSystem
    h::Handle
    nzval::ROCVector{Float64}
    opts::Any
end

"""
    linsolve(F::FactoredSystem, B) -> X

X = A^-1 B with the factor of `F` (the full solve phase, with the refinement `F.opts` asks for).
Differentiable through the `rrule` below with respect to B and to the stored values of A.
"""
function linsolve(F::FactoredSystem, B::ROCMatrix{Float64})
    X = similar(B)
    solve!(F.h, PHASE_SOLVE, X, B, F.nzval, F.opts; solve_mode = 0)
    return X
end

# The adjoint of X = A^-1 B reuses the factor: lambda = A^-T Xbar (a transposed solve), then
# Bbar = lambda and nzvalbar = -lambda X^T scattered onto A's stored pattern.
function ChainRulesCore.rrule(::typeof(linsolve), F::FactoredSystem, B::ROCMatrix{Float64})
    X = linsolve(F, B)
    function linsolve_pullback(Xbar)
        Xb = ROCMatrix{Float64}(unthunk(Xbar))
        lambda = similar(Xb)
        solve!(F.h, PHASE_SOLVE, lambda, Xb, F.nzval, F.opts; solve_mode = 1)
        grad = similar(F.nzval)
        adjoint_scatter!(grad, F.h, lambda, X)
        Fbar = Tangent{FactoredSystem}(; nzval = grad)
        return NoTangent(), Fbar, lambda
    end
    return X, linsolve_pullback
end

end # module
