#=
    SDPX <-> MI355X dense Newton-step provider (optional extension).

    Binds libsdpx_mi355x.so (C-ABI: include/sdpx_mi355x.h) with ccall and
    registers it behind SDPX's two seams:

      * LA provider    -- la_provider_descriptor / instantiate hook
                          (src/la_backend.jl:11-22, 264-276, 508-559);
      * KKT backend    -- a KKTBackend subtype whose factorize!/solve!/
                          solve_direction!/refine! methods are the device
                          calls (src/kkt_backend.jl:24-59, 211-530).

    The dense Newton step also calls residuals, Schur assembly, right-hand
    sides, direction recovery, line search and the update directly
    (src/step.jl:665-936, src/solver/interior_point.jl:1962-2306); one-line
    dispatch hooks at those call sites (step.jl:665, 678-682, 700-706, 847-865,
    934-936; interior_point.jl:1962, 2221) route them to the methods at the end
    of this file when `ws.backend isa MI355XDenseBackend`.

    The batched route (sdpx_batch_*) and the differentiable layer with its
    ChainRulesCore.rrule (the adjoint, sdpx_batch_adjoint) have no counterpart
    in SDPX; they are bound below as well.

    One device handle per SDPX Workspace, called sequentially (the reference's
    one-provider-per-workspace rule, ext/SDPXMultiFloatLinearAlgebraExt.jl:14-20).
    Device buffers are owned by the handle; every host array passed here is
    owned by SDPX and only read or written during the call.
=#
module SDPXMI355XExt

using SDPX
using LinearAlgebra
using Libdl
using ChainRulesCore

const LIB = Ref{Ptr{Cvoid}}(C_NULL)

function _lib()
    if LIB[] == C_NULL
        path = get(ENV, "SDPX_MI355X_LIB",
                   joinpath(@__DIR__, "..", "sdpx_mi355x", "libsdpx_mi355x.so"))
        LIB[] = Libdl.dlopen(path)
    end
    return LIB[]
end

_sym(name::Symbol) = Libdl.dlsym(_lib(), name)

function _check(status::Cint, what::AbstractString)
    status < 0 || return Int(status)
    msg = unsafe_string(ccall(_sym(:sdpx_last_error), Cstring, ()))
    throw(ErrorException("$(what) failed ($(status)): $(msg)"))
end

function _device_count()
    n = Ref{Cint}(0)
    try
        ccall(_sym(:sdpx_device_count), Cint, (Ref{Cint},), n)
    catch
        return 0
    end
    return Int(n[])
end

# ---------------------------------------------------------------------------
# Device handle (replaces the Workspace's dense buffer inventory)
# ---------------------------------------------------------------------------
mutable struct DenseHandle
    ptr::Ptr{Cvoid}
    L::Int
    m::Int
    k::Vector{Int}
end

function DenseHandle(prob::SDPX.SDPProblem{Float64}; device::Int=0)
    L, m, _, k = prob.dims
    out = Ref{Ptr{Cvoid}}(C_NULL)
    kc = Cint.(k)
    _check(ccall(_sym(:sdpx_dense_create), Cint,
                 (Cint, Cint, Ptr{Cint}, Cint, Ref{Ptr{Cvoid}}),
                 device, L, kc, m, out), "sdpx_dense_create")
    h = DenseHandle(out[], L, m, collect(k))
    finalizer(h) do x
        x.ptr == C_NULL || ccall(_sym(:sdpx_dense_destroy), Cint, (Ptr{Cvoid},), x.ptr)
        x.ptr = C_NULL
    end
    if prob.cons isa SDPX.DenseCons{Float64}
        cons = prob.cons::SDPX.DenseCons{Float64}
        for l in 1:L
            # DenseCons.Av[l] is the k^2 x m column-major panel the ABI expects
            _check(ccall(_sym(:sdpx_dense_set_block), Cint,
                         (Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{Float64}),
                         h.ptr, l - 1, cons.Av[l], prob.C[l]), "sdpx_dense_set_block")
        end
    else
        # SparseCons: every block on the sparse pair route, its SparseBlockCOO passed as stored
        # (2x2 blocks keep packed2 instead of a COO; their COO is packed here the same way)
        cons = prob.cons::SDPX.SparseCons{Float64}
        for l in 1:L
            ids = cons.schur_order[l]
            isempty(ids) && throw(ArgumentError("block $l has no active variable"))
            coo = cons.coo[l]
            if length(coo.ptr) != length(ids) + 1
                coo = SDPX.build_block_coo(collect(cons.Asp[l]), ids, k[l])
            end
            _check(ccall(_sym(:sdpx_dense_set_block_sparse), Cint,
                         (Ptr{Cvoid}, Cint, Cint, Ptr{Cint}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32},
                          Ptr{Float64}, Ptr{Float64}),
                         h.ptr, l - 1, length(ids), Cint.(ids .- 1), coo.ptr, coo.lin, coo.row, coo.col, coo.val,
                         prob.C[l]), "sdpx_dense_set_block_sparse")
        end
    end
    _check(ccall(_sym(:sdpx_dense_set_objective), Cint, (Ptr{Cvoid}, Ptr{Float64}),
                 h.ptr, prob.c), "sdpx_dense_set_objective")
    return h
end

_concat(blocks) = reduce(vcat, (vec(b) for b in blocks); init=Float64[])

# Plan-time tiling choices (no counterpart in SDPX): time the Schur build's
# candidates on the device and keep the fastest; returns the 20 reported values
# (panel tile shape codes, split, super-block side, timings, bisection levels).
function autotune!(h::DenseHandle)
    out = zeros(Float64, 20)
    _check(ccall(_sym(:sdpx_dense_autotune), Cint, (Ptr{Cvoid}, Ptr{Float64}), h.ptr, out),
           "sdpx_dense_autotune")
    return out
end

# Re-applies a saved autotune result without timing it.
function set_tiling!(h::DenseHandle, panel0::Integer, panel1::Integer, split::Integer, group::Integer,
                     ftb_levels::Integer)
    _check(ccall(_sym(:sdpx_dense_set_tiling), Cint, (Ptr{Cvoid}, Cint, Cint, Cint, Cint, Cint),
                 h.ptr, panel0, panel1, split, group, ftb_levels), "sdpx_dense_set_tiling")
end

function set_iterate!(h::DenseHandle, x, X, Y)
    _check(ccall(_sym(:sdpx_dense_set_iterate), Cint,
                 (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                 h.ptr, x, _concat(X), _concat(Y)), "sdpx_dense_set_iterate")
end

function get_iterate!(h::DenseHandle, x, X, Y)
    flatX = similar(x, sum(abs2, h.k))
    flatY = similar(flatX)
    _check(ccall(_sym(:sdpx_dense_get_iterate), Cint,
                 (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                 h.ptr, x, flatX, flatY), "sdpx_dense_get_iterate")
    offset = 0
    for l in 1:h.L
        n = h.k[l]^2
        copyto!(X[l], view(flatX, offset+1:offset+n))
        copyto!(Y[l], view(flatY, offset+1:offset+n))
        offset += n
    end
    return x, X, Y
end

# ---------------------------------------------------------------------------
# LA provider seam (src/la_backend.jl:11-22, 283-506; types/backends.jl:26-48)
# ---------------------------------------------------------------------------
# Operation symbols as the planner checks them (la_backend.jl:24-31, 80-97); the
# capability facts below are what the device path implements for Float64.
const _CAPABILITIES = (:chol, :cholesky_factor!, :solve, :multi_rhs, :trsm, :trsv_lower, :trsv_transpose,
                       :syrk, :mul_owned, :dot, :norminf)

const _CAPABILITY_MODEL = SDPX.LAProviderCapabilities(
    cholesky=true,
    factor_solve=true,
    multi_rhs=true,
    iterative_refinement=true,
    mixed_precision_residual=true,   # FP32 / BF16x3 Schur assembly refined against the FP64 operator
    threading=false,                 # one device stream per handle; insensitive to BLAS threads
    dot=true,
    norminf=true,
    mul_owned=true,
    syrk=true,
    triangular_solve=true,
)

function SDPX.la_provider_descriptor(::Type{Float64}, ::Int=1)
    ndev = _device_count()
    ndev > 0 || return (available=false, provider=:none, capabilities=(),
                        capability_model=SDPX.LAProviderCapabilities(), reason=:no_device)
    return (available=true, provider=:mi355x, capabilities=_CAPABILITIES,
            capability_model=_CAPABILITY_MODEL)
end

"""Provider payload of a planned :mi355x LA backend: the device it runs on.  The
dense handle itself is created with the Workspace (`DenseHandle(prob)`), because
it owns the problem's constraint panel in HBM."""
struct MI355XProvider
    device::Int
end

SDPX.la_provider_capability_model(::MI355XProvider) = _CAPABILITY_MODEL

"""Setup hook of the planner branch for `requested === :mi355x` (INTEGRATION.md,
modelled on instantiate_multifloat_la_backend, la_backend.jl:264-276)."""
function instantiate_mi355x_la_backend(::Type{Float64}, config::SDPX.LABackendConfiguration, threads::Int=1)
    config.provider === :mi355x || return nothing
    _device_count() > 0 || return nothing
    return MI355XProvider(0)
end
instantiate_mi355x_la_backend(::Type, ::SDPX.LABackendConfiguration, ::Int=1) = nothing

"""The instantiated :mi355x LA backend.  On the dense Schur route its operations run inside
MI355XDenseBackend on the device (ws.S, its factor and the directions never leave HBM); the
host-buffer la_* methods below serve every other caller, each computed on the device through
the sdpx_la_* entry points (operands uploaded, result written back, la_backend.jl:1529-1769)."""
mutable struct MI355XLABackend <: SDPX.AbstractLABackend
    arithmetic::Symbol
    provider::MI355XProvider
    handle::Ptr{Cvoid}
end
MI355XLABackend(arithmetic::Symbol, provider::MI355XProvider) = MI355XLABackend(arithmetic, provider, C_NULL)

SDPX.la_backend_capabilities(b::MI355XLABackend) = _CAPABILITY_MODEL

function _la(b::MI355XLABackend)
    if b.handle == C_NULL
        out = Ref{Ptr{Cvoid}}(C_NULL)
        _check(ccall(_sym(:sdpx_la_create), Cint, (Cint, Ref{Ptr{Cvoid}}), b.provider.device, out), "sdpx_la_create")
        b.handle = out[]
        finalizer(b) do x
            x.handle == C_NULL || ccall(_sym(:sdpx_la_destroy), Cint, (Ptr{Cvoid},), x.handle)
            x.handle = C_NULL
        end
    end
    return b.handle
end

# Strided column-major Float64 storage: (pointer, leading dimension)
function _colmajor(A::StridedMatrix{Float64})
    stride(A, 1) == 1 || throw(ArgumentError("MI355X provider: column-major operands with unit row stride only"))
    return pointer(A), max(stride(A, 2), size(A, 1), 1)
end
_colmajor(A::StridedVector{Float64}) = (stride(A, 1) == 1 ? pointer(A) :
    throw(ArgumentError("MI355X provider: contiguous vectors only")), max(length(A), 1))

function SDPX.la_chol!(b::MI355XLABackend, A::StridedMatrix{Float64})
    p, ld = _colmajor(A)
    status = GC.@preserve A _check(ccall(_sym(:sdpx_la_chol), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Clonglong),
                                        _la(b), size(A, 1), p, ld), "sdpx_la_chol")
    return status == 0
end

SDPX.la_cholesky_factor!(b::MI355XLABackend, A::StridedMatrix{Float64}) =
    SDPX.la_chol!(b, A) ? SDPX.ProviderLACholeskyFactor(b.provider, A) : nothing

function SDPX.la_trsm!(b::MI355XLABackend, L::StridedMatrix{Float64}, X::StridedVecOrMat{Float64})
    pl, ldl = _colmajor(L)
    px, ldx = X isa StridedVector ? (_colmajor(X)[1], max(length(X), 1)) : _colmajor(X)
    GC.@preserve L X _check(ccall(_sym(:sdpx_la_trsm), Cint,
                                  (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}, Clonglong, Ptr{Float64}, Clonglong),
                                  _la(b), size(L, 1), size(X, 2), pl, ldl, px, ldx), "sdpx_la_trsm")
    return X
end

function _la_trsv!(b::MI355XLABackend, L::StridedMatrix{Float64}, x::StridedVector{Float64}, trans::Integer)
    pl, ldl = _colmajor(L)
    px, _ = _colmajor(x)
    GC.@preserve L x _check(ccall(_sym(:sdpx_la_trsv), Cint,
                                  (Ptr{Cvoid}, Cint, Ptr{Float64}, Clonglong, Ptr{Float64}, Cint),
                                  _la(b), size(L, 1), pl, ldl, px, trans), "sdpx_la_trsv")
    return x
end
SDPX.la_trsv_lower!(b::MI355XLABackend, L, x) = _la_trsv!(b, L, x, 0)
SDPX.la_trsv_transpose!(b::MI355XLABackend, L, x) = _la_trsv!(b, L, x, 1)

function SDPX.la_syrk!(b::MI355XLABackend, S::StridedMatrix{Float64}, P::StridedMatrix{Float64}, α, β)
    ps, lds = _colmajor(S)
    pp, ldp = _colmajor(P)
    GC.@preserve S P _check(ccall(_sym(:sdpx_la_syrk), Cint,
                                  (Ptr{Cvoid}, Cint, Cint, Float64, Ptr{Float64}, Clonglong, Float64, Ptr{Float64},
                                   Clonglong),
                                  _la(b), size(S, 1), size(P, 1), Float64(α), pp, ldp, Float64(β), ps, lds),
                            "sdpx_la_syrk")
    return S
end

# op(A): a Transpose / Adjoint wrapper of a real matrix is passed as its parent with the flag set
_op(A::Union{Transpose{Float64,<:StridedMatrix{Float64}},Adjoint{Float64,<:StridedMatrix{Float64}}}) = (parent(A), 1)
_op(A::StridedMatrix{Float64}) = (A, 0)

function SDPX.la_mul_owned!(b::MI355XLABackend, C::StridedMatrix{Float64}, A, B, α=1.0, β=0.0)
    Ap, ta = _op(A)
    Bp, tb = _op(B)
    pa, lda = _colmajor(Ap)
    pb, ldb = _colmajor(Bp)
    pc, ldc = _colmajor(C)
    m, n = size(C)
    k = size(A, 2)
    GC.@preserve Ap Bp C _check(ccall(_sym(:sdpx_la_gemm), Cint,
                                      (Ptr{Cvoid}, Cint, Cint, Cint, Cint, Cint, Float64, Ptr{Float64}, Clonglong,
                                       Ptr{Float64}, Clonglong, Float64, Ptr{Float64}, Clonglong),
                                      _la(b), ta, tb, m, n, k, Float64(α), pa, lda, pb, ldb, Float64(β), pc, ldc),
                                "sdpx_la_gemm")
    return C
end

function SDPX.la_axpby_owned!(b::MI355XLABackend, α, X::StridedArray{Float64}, β, Y::StridedArray{Float64})
    size(X) == size(Y) || throw(DimensionMismatch("la_axpby_owned!: X and Y differ in size"))
    x, y = vec(X), vec(Y)
    GC.@preserve x y _check(ccall(_sym(:sdpx_la_axpby), Cint,
                                  (Ptr{Cvoid}, Clonglong, Float64, Ptr{Float64}, Float64, Ptr{Float64}),
                                  _la(b), length(y), Float64(α), x, Float64(β), y), "sdpx_la_axpby")
    return Y
end

function SDPX.la_dot(b::MI355XLABackend, x::StridedVector{Float64}, y::StridedVector{Float64})
    out = Ref{Float64}(0.0)
    GC.@preserve x y _check(ccall(_sym(:sdpx_la_dot), Cint, (Ptr{Cvoid}, Clonglong, Ptr{Float64}, Ptr{Float64},
                                                             Ref{Float64}),
                                  _la(b), length(x), x, y, out), "sdpx_la_dot")
    return out[]
end

function SDPX.la_norminf(b::MI355XLABackend, x::StridedVector{Float64})
    out = Ref{Float64}(0.0)
    GC.@preserve x _check(ccall(_sym(:sdpx_la_norminf), Cint, (Ptr{Cvoid}, Clonglong, Ptr{Float64}, Ref{Float64}),
                                _la(b), length(x), x, out), "sdpx_la_norminf")
    return out[]
end

# any other operand type (BigFloat, non-strided views) stays fail-closed (la_backend.jl:1695-1745)
for op in (:la_chol!, :la_cholesky_factor!, :la_trsm!, :la_trsv_lower!, :la_trsv_transpose!, :la_syrk!,
           :la_mul_owned!, :la_axpby_owned!, :la_dot, :la_norminf)
    @eval SDPX.$op(::MI355XLABackend, args...) = throw(ArgumentError(
        "MI355X provider: $($(QuoteNode(op))) takes Float64 strided column-major operands only"))
end

# ---------------------------------------------------------------------------
# Mixed precision (kernels/mixed_precision_kkt.jl template, FP64 target)
# ---------------------------------------------------------------------------
"""Schur Gram precision of the handle: 0 FP64, 1 FP32 MFMA, 2 BF16x3 (bf16 hi/lo
split, three products accumulated in FP32)."""
set_precision!(h::DenseHandle, precision::Integer) =
    _check(ccall(_sym(:sdpx_dense_set_precision), Cint, (Ptr{Cvoid}, Cint), h.ptr, precision),
           "sdpx_dense_set_precision")

"""Schur factor precision: 0 FP64, 1 FP32 (FP32 MFMA panels and trailing updates)."""
set_factor_precision!(h::DenseHandle, precision::Integer) =
    _check(ccall(_sym(:sdpx_dense_set_factor_precision), Cint, (Ptr{Cvoid}, Cint), h.ptr, precision),
           "sdpx_dense_set_factor_precision")

"""Precision of the P-hat products feeding the BF16x3 assembly: 0 FP64, 1 FP32 MFMA (bf16
planes written by the product, FP32 copy of the constraint panel kept on the device)."""
set_panel_precision!(h::DenseHandle, precision::Integer) =
    _check(ccall(_sym(:sdpx_dense_set_panel_precision), Cint, (Ptr{Cvoid}, Cint), h.ptr, precision),
           "sdpx_dense_set_panel_precision")

# Condition estimates from the last Schur factor: cond_1(S) (xPOCON), the xTRCON product of
# _triangular_condition_estimate (mixed_precision_kkt.jl:371-392), rcond_1(L), rcond_inf(L)
function factor_condition(h::DenseHandle; which::Integer=3)
    out = zeros(Float64, 4)
    _check(ccall(_sym(:sdpx_dense_factor_condition), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}), h.ptr, which, out),
           "sdpx_dense_factor_condition")
    return (cond_schur=out[1], cond_trcon=out[2], rcond_1=out[3], rcond_inf=out[4])
end

"""rho = r - S dx against the matrix-free FP64 Schur operator -> (||rho||_inf, ||r||_inf)."""
function exact_residual(h::DenseHandle)
    out = zeros(2)
    _check(ccall(_sym(:sdpx_dense_exact_residual), Cint, (Ptr{Cvoid}, Ptr{Float64}), h.ptr, out),
           "sdpx_dense_exact_residual")
    return out[1], out[2]
end

"""Refinement of dx against the FP64 operator with the current (low-precision-assembled)
factor as preconditioner (_refine_mixed_direction!, mixed_precision_kkt.jl:1634-1713)."""
function refine_exact!(h::DenseHandle, abstol::Float64, max_steps::Integer)
    steps = Ref{Cint}(0)
    residual = Ref{Float64}(0.0)
    _check(ccall(_sym(:sdpx_dense_refine_exact), Cint, (Ptr{Cvoid}, Float64, Cint, Ref{Cint}, Ref{Float64}),
                 h.ptr, abstol, max_steps, steps, residual), "sdpx_dense_refine_exact")
    return Int(steps[]), residual[]
end

# ---------------------------------------------------------------------------
# Batched route: many single-block SDPs in lock step (no counterpart in SDPX)
# ---------------------------------------------------------------------------
"""sdpx_batch_options_t, field for field."""
mutable struct BatchOptions
    beta::Float64; gamma::Float64; eps_gap::Float64; eps_primal::Float64; eps_dual::Float64
    min_step::Float64; omega_step::Float64; stall_tolerance::Float64; refine_tol::Float64
    adaptive_sigma_max::Float64
    iter_max::Cint; max_restarts::Cint; max_centering::Cint; stall_iterations::Cint; restart::Cint
    refine_policy::Cint; refine_steps::Cint; refine_max_steps::Cint; adaptive::Cint; step_rule::Cint
    precision::Cint
    cond_limit::Float64; refine_target::Float64; adjoint_mu::Float64
    mixed_max_refine::Cint
    BatchOptions() = new()
end

"""Reference defaults (SolverOptions{Float64}(), types/workspaces.jl:101-228) at tolerance tol,
overridden by the options SDPX passes."""
"""precision 1: FP32 Schur assembly with FP64 refinement (kernels/mixed_precision_kkt.jl template);
a member whose reduced-precision step fails a guard continues in FP64."""
function BatchOptions(opts::SDPX.SolverOptions{Float64}; precision::Integer=0)
    o = BatchOptions()
    _check(ccall(_sym(:sdpx_batch_default_options), Cint, (Ref{BatchOptions}, Float64), o, opts.ϵ_gap),
           "sdpx_batch_default_options")
    o.beta, o.gamma = opts.β, opts.γ
    o.eps_gap, o.eps_primal, o.eps_dual = opts.ϵ_gap, opts.ϵ_primal, opts.ϵ_dual
    o.min_step, o.omega_step = opts.min_step, opts.omega_step
    o.iter_max, o.refine_steps, o.refine_max_steps = opts.iter_max, opts.refine_steps, opts.refine_max_steps
    o.refine_policy = opts.refine_policy === :fixed ? 1 : 0
    o.adaptive = opts.parameter_strategy === :adaptive ? 1 : 0
    o.adaptive_sigma_max = opts.adaptive_sigma_max
    o.precision = precision
    return o
end

mutable struct BatchHandle
    ptr::Ptr{Cvoid}
    k::Vector{Int}
    m::Vector{Int}
end

function BatchHandle(k::AbstractVector{<:Integer}, m::AbstractVector{<:Integer}; device::Int=0)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    _check(ccall(_sym(:sdpx_batch_create), Cint, (Cint, Cint, Ptr{Cint}, Ptr{Cint}, Ref{Ptr{Cvoid}}),
                 device, length(k), Cint.(k), Cint.(m), out), "sdpx_batch_create")
    h = BatchHandle(out[], collect(Int, k), collect(Int, m))
    finalizer(h) do x
        x.ptr == C_NULL || ccall(_sym(:sdpx_batch_destroy), Cint, (Ptr{Cvoid},), x.ptr)
        x.ptr = C_NULL
    end
    return h
end

"""Members given as SDPX problems with one dense PSD block each (DenseCons.Av[1] as the panel)."""
function set_data!(h::BatchHandle, probs::AbstractVector{<:SDPX.SDPProblem{Float64}})
    Av = reduce(vcat, (vec((p.cons::SDPX.DenseCons{Float64}).Av[1]) for p in probs))
    C = reduce(vcat, (vec(p.C[1]) for p in probs))
    c = reduce(vcat, (p.c for p in probs))
    _check(ccall(_sym(:sdpx_batch_set_dense), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                 h.ptr, Av, C, c), "sdpx_batch_set_dense")
end

"""Members' sparse constraints as SDPX stores them: one SparseBlockCOO per member
(types/constraints.jl:59-109), passed as stored (1-based Int32, both triangles, lin)."""
function set_data!(h::BatchHandle, coos::AbstractVector{<:SDPX.SparseBlockCOO{Float64}}, Cs, cs)
    ptr = reduce(vcat, (c.ptr for c in coos))
    lin = reduce(vcat, (c.lin for c in coos); init=Int32[])
    row = reduce(vcat, (c.row for c in coos); init=Int32[])
    col = reduce(vcat, (c.col for c in coos); init=Int32[])
    val = reduce(vcat, (c.val for c in coos); init=Float64[])
    _check(ccall(_sym(:sdpx_batch_set_block_coo), Cint,
                 (Ptr{Cvoid}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Float64}, Ptr{Float64},
                  Ptr{Float64}),
                 h.ptr, ptr, lin, row, col, val, reduce(vcat, vec.(Cs)), reduce(vcat, cs)),
           "sdpx_batch_set_block_coo")
end

function set_start!(h::BatchHandle, xs, Xs, Ys)
    _check(ccall(_sym(:sdpx_batch_set_start), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                 h.ptr, reduce(vcat, xs), reduce(vcat, vec.(Xs)), reduce(vcat, vec.(Ys))),
           "sdpx_batch_set_start")
end

const BATCH_STATUS = (:Optimal, :IterLimit, :NumericalBreakdown, :Stalled, :InsufficientPrecision,
                      :MaxRestartsExceeded)

"""Solve every member; returns (status, iterations, pobj, dobj, x, X, Y) per member."""
function batch_solve!(h::BatchHandle, o::BatchOptions)
    launched = Ref{Cint}(0)
    _check(ccall(_sym(:sdpx_batch_solve), Cint, (Ptr{Cvoid}, Ref{BatchOptions}, Ref{Cint}), h.ptr, o, launched),
           "sdpx_batch_solve")
    B = length(h.k)
    status, iters = zeros(Cint, B), zeros(Cint, B)
    pobj, dobj = zeros(B), zeros(B)
    x = zeros(sum(h.m))
    X, Y = zeros(sum(abs2, h.k)), zeros(sum(abs2, h.k))
    _check(ccall(_sym(:sdpx_batch_get_result), Cint,
                 (Ptr{Cvoid}, Ptr{Cint}, Ptr{Cint}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                  Ptr{Float64}), h.ptr, status, iters, pobj, dobj, x, X, Y), "sdpx_batch_get_result")
    xs = _split(x, h.m)
    Xs = [reshape(b, k, k) for (b, k) in zip(_split(X, abs2.(h.k)), h.k)]
    Ys = [reshape(b, k, k) for (b, k) in zip(_split(Y, abs2.(h.k)), h.k)]
    return (status=[BATCH_STATUS[s+1] for s in status], iterations=Int.(iters), pobj=pobj, dobj=dobj,
            x=xs, X=Xs, Y=Ys)
end

function _split(v, sizes)
    out = Vector{Vector{Float64}}(undef, length(sizes))
    o = 0
    for (i, n) in enumerate(sizes)
        out[i] = v[o+1:o+n]
        o += n
    end
    return out
end

"""Backward pass of the layer at the solved batch: restore the differentiation checkpoints,
centre them (`iterations` Newton steps at fixed mu) and run the adjoint for cotangents
G (of X*) and hv (of y*).  Returns (dC, db) per member (NaN where no checkpoint)."""
# The optimum's derivative as the mu -> 0 limit of the central-path adjoint: the checkpoint
# re-centred at its own mu (D1) and at ratio * mu (D2), combined as (D2 - ratio D1) / (1 - ratio)
function batch_adjoint!(h::BatchHandle, Gs, hs; iterations::Integer=6, ratio::Float64=0.1)
    B = length(h.k)
    restored = Ref{Cint}(0)
    mu = zeros(B)
    _check(ccall(_sym(:sdpx_batch_restore_checkpoint), Cint, (Ptr{Cvoid}, Ref{Cint}, Ptr{Float64}),
                 h.ptr, restored, mu), "sdpx_batch_restore_checkpoint")
    G, hv = reduce(vcat, vec.(Gs)), reduce(vcat, hs)
    function one_adjoint()
        dC, db = zeros(sum(abs2, h.k)), zeros(sum(h.m))
        _check(ccall(_sym(:sdpx_batch_adjoint), Cint,
                     (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                     h.ptr, G, hv, dC, db), "sdpx_batch_adjoint")
        return dC, db
    end
    _check(ccall(_sym(:sdpx_batch_center), Cint, (Ptr{Cvoid}, Cint), h.ptr, iterations), "sdpx_batch_center")
    dC1, db1 = one_adjoint()
    _check(ccall(_sym(:sdpx_batch_center_at), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}), h.ptr, iterations,
                 ratio .* mu), "sdpx_batch_center_at")
    dC2, db2 = one_adjoint()
    dC = (dC2 .- ratio .* dC1) ./ (1 - ratio)
    db = (db2 .- ratio .* db1) ./ (1 - ratio)
    dCs = [reshape(b, k, k) for (b, k) in zip(_split(dC, abs2.(h.k)), h.k)]
    dbs = _split(db, h.m)
    for p in 1:B
        if isnan(mu[p])
            fill!(dCs[p], NaN)
            fill!(dbs[p], NaN)
        end
    end
    return dCs, dbs
end

# ---------------------------------------------------------------------------
# Differentiable SDP layer (no counterpart in SDPX: it has no differentiation)
# ---------------------------------------------------------------------------
"""A batch of user-model SDPs  max b'y  s.t.  C - sum_i y_i A_i >= 0  with fixed A
(one PSD block each) whose data (C, b) is the layer's input and whose optimum
(X*, y*) is its output.  In SDPX's core form (ir/lower_sdp.jl:244-300) x = y,
A_core = -A, C_core = -C, c = -b, and X* is the core Y."""
struct SDPLayer
    handle::BatchHandle
    Av::Vector{Matrix{Float64}}   # per member k^2 x m panel of the core A (DenseCons.Av[1])
    x0::Vector{Vector{Float64}}
    options::BatchOptions
end

function SDPLayer(Av::AbstractVector, x0::AbstractVector, opts::SDPX.SolverOptions{Float64}; device::Int=0)
    k = [isqrt(size(a, 1)) for a in Av]
    m = [size(a, 2) for a in Av]
    return SDPLayer(BatchHandle(k, m; device=device), [Matrix{Float64}(a) for a in Av],
                    [Vector{Float64}(v) for v in x0], BatchOptions(opts))
end

"""Forward solve: (X*_p, y*_p) for every member at data (C_p, b_p) of the user's model."""
function (layer::SDPLayer)(Cs::AbstractVector, bs::AbstractVector)
    h = layer.handle
    Av = reduce(vcat, vec.(layer.Av))
    Ccore = reduce(vcat, (vec(-C) for C in Cs))
    ccore = reduce(vcat, (-b for b in bs))
    _check(ccall(_sym(:sdpx_batch_set_dense), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                 h.ptr, Av, Ccore, ccore), "sdpx_batch_set_dense")
    eyes = [Matrix{Float64}(I, k, k) for k in h.k]
    set_start!(h, layer.x0, eyes, eyes)
    res = batch_solve!(h, layer.options)
    return res.Y, res.x   # X* (the core Y) and y* (the core x)
end

"""The layer's reverse rule: for cotangents (Xbar_p, ybar_p) of (X*_p, y*_p) the adjoint
of L = sum_p <Xbar_p, X*_p> + ybar_p'y*_p gives dL/dC_p and dL/db_p, with the solve's
Schur Cholesky refactored at the re-centred differentiation checkpoint (sdpx_batch_adjoint)."""
function ChainRulesCore.rrule(layer::SDPLayer, Cs::AbstractVector, bs::AbstractVector)
    Xs, ys = layer(Cs, bs)
    function sdp_layer_pullback(ȳ)
        Xbar, ybar = ChainRulesCore.unthunk(ȳ)
        G = [Xbar isa ChainRulesCore.AbstractZero ? zeros(size(X)) : Matrix{Float64}(Xbar[p])
             for (p, X) in enumerate(Xs)]
        hv = [ybar isa ChainRulesCore.AbstractZero ? zeros(length(y)) : Vector{Float64}(ybar[p])
              for (p, y) in enumerate(ys)]
        dCs, dbs = batch_adjoint!(layer.handle, G, hv)
        return ChainRulesCore.NoTangent(), dCs, dbs
    end
    return (Xs, ys), sdp_layer_pullback
end

# ---------------------------------------------------------------------------
# KKT backend seam
# ---------------------------------------------------------------------------
"""Dense Cholesky KKT backend whose Schur matrix, factor and directions live on
one MI355X; selected by the execution plan when the provider is :mi355x."""
mutable struct MI355XDenseBackend <: SDPX.KKTBackend
    handle::DenseHandle
end

SDPX.backend_name(::MI355XDenseBackend) = :mi355x_dense_cholesky

"""Block-arrow KKT backend (the :block_arrow route, factor_arrow_kkt! / solve_arrow_kkt!): the
same handle with its factor and solves switched to the arrow elimination over the partition of
the workspace's ArrowWorkspace (local_ids per block, global_ids)."""
function arrow_backend(h::DenseHandle, arrow::SDPX.ArrowWorkspace{Float64})
    owner = fill(Cint(-1), h.m)
    for (l, ids) in pairs(arrow.local_ids), i in ids
        owner[i] = Cint(l - 1)
    end
    _check(ccall(_sym(:sdpx_dense_set_arrow), Cint, (Ptr{Cvoid}, Ptr{Cint}), h.ptr, owner),
           "sdpx_dense_set_arrow")
    return MI355XDenseBackend(h)
end

function SDPX.factorize!(b::MI355XDenseBackend, ws::SDPX.Workspace{Float64},
                         prob::SDPX.SDPProblem{Float64}, opts::SDPX.SolverOptions{Float64})
    attempts = Ref{Cint}(0)
    status = _check(ccall(_sym(:sdpx_dense_schur_factor), Cint, (Ptr{Cvoid}, Ref{Cint}),
                          b.handle.ptr, attempts), "sdpx_dense_schur_factor")
    return (ok=status == 0, reg_attempts=Int(attempts[]), q_pivoted=false,
            q_rank_deficient=false, equality_solver=:normal_equations,
            phase_times=(schur_copy=0.0, schur_factorization=0.0,
                         constraint_triangular_solve=0.0, equality_gram=0.0,
                         equality_factorization=0.0))
end

function SDPX.solve!(b::MI355XDenseBackend, ws::SDPX.Workspace{Float64}, n::Int,
                     r::AbstractVector{Float64}, p_rhs, dx_out, dy_out)
    _check(ccall(_sym(:sdpx_dense_set_rhs), Cint, (Ptr{Cvoid}, Ptr{Float64}), b.handle.ptr, r),
           "sdpx_dense_set_rhs")
    _check(ccall(_sym(:sdpx_dense_solve), Cint, (Ptr{Cvoid},), b.handle.ptr), "sdpx_dense_solve")
    _check(ccall(_sym(:sdpx_dense_get_direction), Cint,
                 (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                 b.handle.ptr, dx_out, C_NULL, C_NULL), "sdpx_dense_get_direction")
    return dx_out, dy_out
end

function SDPX.solve_direction!(b::MI355XDenseBackend, ws::SDPX.Workspace{Float64},
                               prob::SDPX.SDPProblem{Float64}, opts::SDPX.SolverOptions{Float64},
                               r::AbstractVector{Float64})
    # r already lives on the device (produced by the right-hand-side hook)
    _check(ccall(_sym(:sdpx_dense_solve), Cint, (Ptr{Cvoid},), b.handle.ptr), "sdpx_dense_solve")
    return true
end

function SDPX.refine!(b::MI355XDenseBackend, ws::SDPX.Workspace{Float64},
                      prob::SDPX.SDPProblem{Float64}, opts::SDPX.SolverOptions{Float64},
                      r::AbstractVector{Float64})
    out = zeros(2)
    _check(ccall(_sym(:sdpx_dense_direction_residual), Cint, (Ptr{Cvoid}, Ptr{Float64}),
                 b.handle.ptr, out), "sdpx_dense_direction_residual")
    fixed = opts.refine_policy === :fixed
    cap = fixed ? opts.refine_steps : opts.refine_max_steps
    rel = opts.refine_tol > 0 ? opts.refine_tol : SDPX.REFINE_DEFAULT_TOL_ULPS * eps(Float64)
    steps = Ref{Cint}(0)
    residual = Ref{Float64}(0.0)
    _check(ccall(_sym(:sdpx_dense_refine), Cint,
                 (Ptr{Cvoid}, Float64, Cint, Cint, Ref{Cint}, Ref{Float64}),
                 b.handle.ptr, fixed ? 0.0 : rel * max(out[2], 1.0), cap, fixed, steps, residual),
           "sdpx_dense_refine")
    return (Int(steps[]), residual[])
end

# ---------------------------------------------------------------------------
# Newton-step hooks (called from step.jl / interior_point.jl when
# ws.backend isa MI355XDenseBackend)
# ---------------------------------------------------------------------------
function mi355x_compute_residuals!(b::MI355XDenseBackend, μ, opts; factor::Bool=true)
    out = zeros(3)
    status = _check(ccall(_sym(:sdpx_dense_residuals), Cint,
                          (Ptr{Cvoid}, Ptr{Float64}, Cint, Cint, Ptr{Float64}),
                          b.handle.ptr, μ, opts.parameter_strategy === :adaptive, factor, out),
                    "sdpx_dense_residuals")
    return out[1], out[2], status == 0
end

mi355x_schur_build!(b::MI355XDenseBackend) =
    _check(ccall(_sym(:sdpx_dense_schur_build), Cint, (Ptr{Cvoid},), b.handle.ptr), "sdpx_dense_schur_build")

mi355x_predictor_rhs!(b::MI355XDenseBackend) =
    _check(ccall(_sym(:sdpx_dense_predictor_rhs), Cint, (Ptr{Cvoid},), b.handle.ptr), "sdpx_dense_predictor_rhs")

function mi355x_corrector_rhs!(b::MI355XDenseBackend, mehrotra::Bool, σ, μ_global, block_local::Bool, μ, β)
    _check(ccall(_sym(:sdpx_dense_corrector_rhs), Cint,
                 (Ptr{Cvoid}, Cint, Float64, Float64, Cint, Ptr{Float64}, Float64),
                 b.handle.ptr, mehrotra, σ, μ_global, block_local, μ, β), "sdpx_dense_corrector_rhs")
end

mi355x_direction_blocks!(b::MI355XDenseBackend) =
    _check(ccall(_sym(:sdpx_dense_direction_blocks), Cint, (Ptr{Cvoid},), b.handle.ptr),
           "sdpx_dense_direction_blocks")

function mi355x_line_search!(b::MI355XDenseBackend, rule::Symbol, pf, df, γ, min_step, ratio)
    tX = Ref{Float64}(1.0)
    tY = Ref{Float64}(1.0)
    if rule === :fraction_to_boundary
        _check(ccall(_sym(:sdpx_dense_fraction_to_boundary), Cint,
                     (Ptr{Cvoid}, Float64, Float64, Ref{Float64}, Ref{Float64}),
                     b.handle.ptr, pf, df, tX, tY), "sdpx_dense_fraction_to_boundary")
    else
        _check(ccall(_sym(:sdpx_dense_backtrack), Cint,
                     (Ptr{Cvoid}, Float64, Float64, Float64, Ref{Float64}, Ref{Float64}),
                     b.handle.ptr, γ, min_step, ratio, tX, tY), "sdpx_dense_backtrack")
    end
    return tX[], tY[]
end

function mi355x_update!(b::MI355XDenseBackend, tX, tY)
    out = zeros(2 + b.handle.L)
    _check(ccall(_sym(:sdpx_dense_update), Cint, (Ptr{Cvoid}, Float64, Float64, Ptr{Float64}),
                 b.handle.ptr, tX, tY, out), "sdpx_dense_update")
    return out[1], out[2] != 0.0, out[3:end]
end

end # module
