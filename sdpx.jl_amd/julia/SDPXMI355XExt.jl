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

    One device handle per SDPX Workspace, called sequentially (the reference's
    one-provider-per-workspace rule, ext/SDPXMultiFloatLinearAlgebraExt.jl:14-20).
    Device buffers are owned by the handle; every host array passed here is
    owned by SDPX and only read or written during the call.
=#
module SDPXMI355XExt

using SDPX
using LinearAlgebra
using Libdl

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
    cons = prob.cons::SDPX.DenseCons{Float64}
    for l in 1:L
        # DenseCons.Av[l] is the k^2 x m column-major panel the ABI expects
        _check(ccall(_sym(:sdpx_dense_set_block), Cint,
                     (Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{Float64}),
                     h.ptr, l - 1, cons.Av[l], prob.C[l]), "sdpx_dense_set_block")
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
# LA provider seam
# ---------------------------------------------------------------------------
const _CAPABILITIES = (:chol, :cholesky_factor!, :solve, :trsv_lower, :trsv_transpose,
                       :syrk, :mul_owned, :dot, :multi_rhs)

function SDPX.la_provider_descriptor(::Type{Float64}, ::Int=1)
    available = _device_count() > 0
    return (available=available, provider=:mi355x,
            capabilities=available ? _CAPABILITIES : ())
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
