"""
    LowRankIntegratorsHIP

Runs the per-step hot path of LowRankIntegrators.jl on an AMD MI355X through
`liblri_hip.so` (declared in `include/lri.h`), behind the reference's own
operator API: wrap any of its algorithms in [`HIPAlgorithm`](@ref) and call
the reference's unmodified `solve`,

```julia
using LowRankIntegrators, LowRankIntegratorsHIP
alg = HIPAlgorithm(ProjectorSplitting(PrimalLieTrotter()); precision = :mixed, snapshot_dtype = :bf16)
sol = LowRankIntegrators.solve(MatrixDataProblem(snapshots, u0), alg)
```

`solve` (primitives.jl:68-80) dispatches to the `init`, `step!` and
`update_sol!` methods below: `init` builds a device context in place of the
reference's `alg_cache` (projector_splitting.jl:87-105, unconventional.jl:86-107,
rank_adaptive_unconventional.jl:106-131, greedy_integrator.jl:24-39), `step!`
replaces the bodies at projector_splitting.jl:191-211, unconventional.jl:159-164,
rank_adaptive_unconventional.jl:171-180 and greedy_integrator.jl:106-111 by one
`ccall` each, and `update_sol!` keeps the per-step copies of primitives.jl:82-90
in a device trajectory that is copied out once, after the last step (the same
`(tf - t)/T > 1e-8` rule as `solve` decides which step is the last).

Matrix ODEs (`MatrixDEProblem`) take their right-hand side as an operator
descriptor, [`LyapunovOperator`](@ref) or [`QuadraticOperator`](@ref), in place
of a closure, which cannot run on the device.  The reverse of the BUG, primal
KSL and rank-adaptive data steps sits behind a ChainRulesCore `rrule` of
[`final_state`](@ref).

Every call returns a status; a nonzero one is raised as `ArgumentError` (bad
argument: the reference raises `AssertionError`/`DimensionMismatch` there) or
as [`LRIError`](@ref) with the library's message.
"""
module LowRankIntegratorsHIP

using LinearAlgebra
using ChainRulesCore
using LowRankIntegrators
using LowRankIntegrators: AbstractDLRAlgorithm, AbstractDLRAlgorithm_Cache, DLRIntegrator, DLRSolution, init_sol
using LowRankArithmetic: SVDLikeRepresentation, TwoFactorRepresentation
import LowRankIntegrators: init, step!, alg_cache, update_sol!

export HIPAlgorithm, HIPContext, LRIError, LyapunovOperator, QuadraticOperator, final_state, solve_gradient,
       solve_batch, autotune!

# ---------------------------------------------------------------------------- library

"The shared library (override with the environment variable `LRI_HIP_LIB`)."
const libhip = get(ENV, "LRI_HIP_LIB", normpath(joinpath(@__DIR__, "..", "lowrankintegrators.jl_amd", "liblri_hip.so")))

# include/lri.h
const LRI_OK = Cint(0)
const LRI_EINVAL = Cint(-1)
const LRI_F64, LRI_F32, LRI_BF16 = Cint(0), Cint(1), Cint(2)
const LRI_PREC_F64, LRI_PREC_MIXED = Cint(0), Cint(1)
const LRI_PRIMAL, LRI_DUAL, LRI_STRANG = Cint(0), Cint(1), Cint(2)
const LRI_MAX_RANK = 64
const DTYPES = Dict(:f64 => LRI_F64, :f32 => LRI_F32, :bf16 => LRI_BF16)

"A failure reported by `liblri_hip.so` (status `LRI_E*`, message of `lri_last_error`)."
struct LRIError <: Exception
    status::Cint
    msg::String
end
Base.showerror(io::IO, e::LRIError) = print(io, "liblri_hip: ", e.msg, " (status ", e.status, ")")

function check(ctx::Ptr{Cvoid}, st::Cint)
    st == LRI_OK && return nothing
    msg = unsafe_string(ccall((:lri_last_error, libhip), Cstring, (Ptr{Cvoid},), ctx))
    st == LRI_EINVAL && throw(ArgumentError(msg))
    throw(LRIError(st, msg))
end

# ---------------------------------------------------------------------------- context

"""
    HIPContext(n, m, rcap; device = 0)

One problem size (n x m data, ranks up to `rcap`) on one GPU: `lri_create`.
The context owns its device buffers; a finalizer calls `lri_destroy`.
"""
mutable struct HIPContext
    ptr::Ptr{Cvoid}
    n::Int
    m::Int
    rcap::Int
    resident::Bool             # lri_set_resident_eigenbasis is on
    function HIPContext(n::Integer, m::Integer, rcap::Integer; device::Integer = 0)
        out = Ref{Ptr{Cvoid}}(C_NULL)
        st = ccall((:lri_create, libhip), Cint, (Ref{Ptr{Cvoid}}, Cint, Int64, Int64, Int64),
                   out, device, n, m, rcap)
        check(C_NULL, st)
        c = new(out[], Int(n), Int(m), Int(rcap), false)
        finalizer(destroy!, c)
        return c
    end
end

function destroy!(c::HIPContext)
    if c.ptr != C_NULL
        ccall((:lri_destroy, libhip), Cint, (Ptr{Cvoid},), c.ptr)
        c.ptr = C_NULL
    end
    return nothing
end

"A context owned by an `lri_batch` (`lri_batch_context`): the context API without ownership."
struct MemberContext
    ptr::Ptr{Cvoid}
    n::Int
    m::Int
    rcap::Int
end

const AnyContext = Union{HIPContext,MemberContext}

set_precision!(c::AnyContext, prec::Symbol) =
    check(c.ptr, ccall((:lri_set_precision, libhip), Cint, (Ptr{Cvoid}, Cint), c.ptr,
                       prec === :mixed ? LRI_PREC_MIXED : LRI_PREC_F64))

"`lri_set_state`: the `u = deepcopy(prob.u0)` of `init` (projector_splitting.jl:110)."
function set_state!(c::AnyContext, U::AbstractMatrix, S::AbstractMatrix, V::AbstractMatrix)
    Uc, Sc, Vc = Matrix{Float64}(U), Matrix{Float64}(S), Matrix{Float64}(V)
    GC.@preserve Uc Sc Vc check(c.ptr, ccall((:lri_set_state, libhip), Cint,
        (Ptr{Cvoid}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64),
        c.ptr, size(Sc, 1), Uc, size(Uc, 1), Sc, size(Sc, 1), Vc, size(Vc, 1)))
end
set_state!(c::AnyContext, u::SVDLikeRepresentation) = set_state!(c, u.U, u.S, u.V)
# a two-factor state U Z' keeps Z in the V slot (S unused)
set_state!(c::AnyContext, u::TwoFactorRepresentation) = set_state!(c, u.U, Matrix{Float64}(I, size(u.U, 2), size(u.U, 2)), u.Z)

rank_of(c::AnyContext) = (r = Ref{Int64}(0); check(c.ptr, ccall((:lri_get_rank, libhip), Cint,
                                                                 (Ptr{Cvoid}, Ref{Int64}), c.ptr, r)); Int(r[]))

"`lri_get_state` as a tuple `(U, S, V)` (synchronising)."
function get_factors(c::AnyContext)
    r = rank_of(c)
    U, S, V = zeros(c.n, r), zeros(r, r), zeros(c.m, r)
    rout = Ref{Int64}(0)
    check(c.ptr, ccall((:lri_get_state, libhip), Cint,
        (Ptr{Cvoid}, Ref{Int64}, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64),
        c.ptr, rout, U, c.n, S, r, V, c.m))
    return U, S, V
end
get_state(c::AnyContext) = SVDLikeRepresentation(get_factors(c)...)
get_state_twofactor(c::HIPContext) = (f = get_factors(c); TwoFactorRepresentation(f[1], f[3]))

# ---------------------------------------------------------------------------- data

# BF16 bits of a Float32 (round to nearest even), for BF16 snapshot slots
bf16_bits(x::Float32) = (u = reinterpret(UInt32, x); UInt16((u + 0x00007fff + ((u >> 16) & 0x1)) >> 16))

function encode(A::AbstractMatrix, dtype::Symbol)
    dtype === :f64 && return Matrix{Float64}(A)
    dtype === :f32 && return Matrix{Float32}(A)
    dtype === :bf16 && return bf16_bits.(Matrix{Float32}(A))
    throw(ArgumentError("snapshot dtype must be :f64, :f32 or :bf16"))
end

reserve_snapshots!(c::HIPContext, count::Integer, dtype::Symbol) =
    check(c.ptr, ccall((:lri_snapshots_reserve, libhip), Cint, (Ptr{Cvoid}, Int64, Cint), c.ptr, count, DTYPES[dtype]))

"`lri_snapshot_upload` of A into slot k (0-based): the data of `update_data!` (data_integrator.jl:22-28)."
function upload_snapshot!(c::HIPContext, k::Integer, A::AbstractMatrix, dtype::Symbol)
    B = encode(A, dtype)
    GC.@preserve B check(c.ptr, ccall((:lri_snapshot_upload, libhip), Cint,
                                      (Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64), c.ptr, k, B, size(B, 1)))
end

# ---------------------------------------------------------------------------- matrix-ODE operators

"""
    LyapunovOperator(A::SymTridiagonal, B::SymTridiagonal, G, H; periodic_A = false, periodic_B = false)

F(Y) = A Y + Y B' + G H' (G n x q, H m x q, q <= 4): the device form of a
`MatrixDEProblem`'s right-hand side (`lri_set_operator_lyapunov`).  With
`periodic_*` the off-diagonal vector has n entries, the last one coupling the
last and the first row.  Calling it evaluates F on a dense matrix.
"""
struct LyapunovOperator
    ad::Vector{Float64}
    ae::Vector{Float64}
    aper::Bool
    bd::Vector{Float64}
    be::Vector{Float64}
    bper::Bool
    G::Matrix{Float64}
    H::Matrix{Float64}
end
function LyapunovOperator(A::SymTridiagonal, B::SymTridiagonal, G = zeros(size(A, 1), 0), H = zeros(size(B, 1), 0);
                          periodic_A::Bool = false, periodic_B::Bool = false, ae_wrap = nothing, be_wrap = nothing)
    ae = periodic_A ? vcat(A.ev, something(ae_wrap, A.ev[1])) : Vector{Float64}(A.ev)
    be = periodic_B ? vcat(B.ev, something(be_wrap, B.ev[1])) : Vector{Float64}(B.ev)
    return LyapunovOperator(Vector{Float64}(A.dv), ae, periodic_A, Vector{Float64}(B.dv), be, periodic_B,
                            Matrix{Float64}(reshape(G, size(A, 1), :)), Matrix{Float64}(reshape(H, size(B, 1), :)))
end

function tridiag_apply(d, e, per, Y)
    n = length(d)
    Z = d .* Y
    for i in 1:n-1
        Z[i, :] .+= e[i] .* Y[i+1, :]
        Z[i+1, :] .+= e[i] .* Y[i, :]
    end
    if per
        Z[n, :] .+= e[n] .* Y[1, :]
        Z[1, :] .+= e[n] .* Y[n, :]
    end
    return Z
end
(F::LyapunovOperator)(Y::AbstractMatrix, t = 0.0) =
    tridiag_apply(F.ad, F.ae, F.aper, Y) + permutedims(tridiag_apply(F.bd, F.be, F.bper, permutedims(Y))) + F.G * F.H'

"""
    QuadraticOperator(L::LyapunovOperator, dl, dd, du; periodic = false)

F(Y) = A Y + Y B' + G H' - (D Y) .* Y with D = tridiag(dl, dd, du) on the n rows
(`lri_set_operator_quadratic`): viscous Burgers with one parameter sample per
column, `Δ*ρ - (∇*ρ) .* ρ` (test/data_agnostic_approximation.jl:31-33).
"""
struct QuadraticOperator
    L::LyapunovOperator
    dl::Vector{Float64}
    dd::Vector{Float64}
    du::Vector{Float64}
    periodic::Bool
end
QuadraticOperator(L::LyapunovOperator, dl, dd, du; periodic::Bool = false) =
    QuadraticOperator(L, Vector{Float64}(dl), Vector{Float64}(dd), Vector{Float64}(du), periodic)
function (F::QuadraticOperator)(Y::AbstractMatrix, t = 0.0)
    n = size(Y, 1)
    DY = F.dd .* Y
    for i in 1:n
        (i > 1 || F.periodic) && (DY[i, :] .+= F.dl[i] .* Y[i > 1 ? i - 1 : n, :])
        (i < n || F.periodic) && (DY[i, :] .+= F.du[i] .* Y[i < n ? i + 1 : 1, :])
    end
    return F.L(Y, t) - DY .* Y
end

function set_operator!(c::AnyContext, F::LyapunovOperator)
    q = size(F.G, 2)
    G, H = F.G, F.H
    check(c.ptr, GC.@preserve F G H ccall((:lri_set_operator_lyapunov, libhip), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Float64}, Ptr{Float64}, Cint, Int64,
         Ptr{Float64}, Int64, Ptr{Float64}, Int64),
        c.ptr, F.ad, F.ae, Cint(F.aper), F.bd, F.be, Cint(F.bper), q,
        q > 0 ? pointer(G) : Ptr{Float64}(C_NULL), size(G, 1), q > 0 ? pointer(H) : Ptr{Float64}(C_NULL), size(H, 1)))
end
function set_operator!(c::AnyContext, F::QuadraticOperator)
    set_operator!(c, F.L)
    check(c.ptr, GC.@preserve F ccall((:lri_set_operator_quadratic, libhip), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cint),
        c.ptr, F.dl, F.dd, F.du, Cint(F.periodic)))
end

# Constant coefficients on both sides of a linear operator: the state stays in the
# operator's analytic eigenbasis (lri_set_resident_eigenbasis), so a step needs no
# n x n transforms.
is_constant(d, e) = all(==(d[1]), d) && (isempty(e) || all(==(e[1]), e))
use_eigenbasis!(c::AnyContext, F::QuadraticOperator) = nothing
function use_eigenbasis!(c::AnyContext, F::LyapunovOperator)
    (is_constant(F.ad, F.ae) && is_constant(F.bd, F.be)) || return nothing
    check(c.ptr, ccall((:lri_set_resident_eigenbasis, libhip), Cint, (Ptr{Cvoid}, Cint), c.ptr, Cint(1)))
    c isa HIPContext && (c.resident = true)
    return nothing
end

# Gershgorin bound of the operator's linear part, for the default RK4 sub-step count
gersh(d, e, per) = maximum(abs(d[i]) + (i > 1 || per ? abs(e[i > 1 ? i - 1 : end]) : 0.0) +
                           (i < length(d) || per ? abs(e[min(i, length(e))]) : 0.0) for i in eachindex(d))
spectral_bound(F::LyapunovOperator) = gersh(F.ad, F.ae, F.aper) + gersh(F.bd, F.be, F.bper)
spectral_bound(F::QuadraticOperator, ymax) =
    spectral_bound(F.L) + 2 * ymax * maximum(abs.(F.dl) .+ abs.(F.dd) .+ abs.(F.du))
"Fewest classical RK4 steps with h * rho <= 2.5 (RK4's real stability interval is [-2.78, 0])."
rk4_substeps(rho, dt) = max(1, ceil(Int, abs(dt) * rho / 2.5))

# ---------------------------------------------------------------------------- the algorithm wrapper

"""
    HIPAlgorithm(alg; precision = :f64, snapshot_dtype = :f64, substeps = nothing, device = 0)

Runs `alg` (`ProjectorSplitting(order)`, `UnconventionalAlgorithm()`,
`RankAdaptiveUnconventionalAlgorithm(tol; rmax)` or `GreedyIntegrator()`) on
the GPU.  `precision = :mixed` runs the n x m x r products on BF16 (BF16
snapshots, hi/lo split) or FP32 (FP32 snapshots) matrix cores; everything else
stays FP64.  `substeps`: classical RK4 sub-steps per step of a matrix ODE (the
reference integrates its sub-problems with adaptive Tsit5); `nothing` picks
the fewest stable ones.
"""
struct HIPAlgorithm{A<:AbstractDLRAlgorithm} <: AbstractDLRAlgorithm
    alg::A
    precision::Symbol
    snapshot_dtype::Symbol
    substeps::Union{Nothing,Int}
    device::Int
end
HIPAlgorithm(alg::AbstractDLRAlgorithm; precision::Symbol = :f64, snapshot_dtype::Symbol = :f64,
             substeps::Union{Nothing,Integer} = nothing, device::Integer = 0) =
    HIPAlgorithm(alg, precision, snapshot_dtype, substeps === nothing ? nothing : Int(substeps), Int(device))

"The device side of a `DLRIntegrator` whose algorithm is a `HIPAlgorithm`."
mutable struct HIPCache <: AbstractDLRAlgorithm_Cache
    ctx::HIPContext
    y::Any                     # the data of a MatrixDataProblem (array or function), else nothing
    array_data::Bool
    kprev::Int                 # array data: the 0-based slot of yprev (alg_cache seeds it with y[1])
    slot::Int                  # function data: the ring slot (of 3) holding y(t)
    rank::Int
    twofactor::Bool
    substeps::Int              # matrix ODE: RK4 sub-steps per step
    quad_ymax::Float64         # quadratic operator: bound on |Y| for the default sub-step count
    tspan::Tuple{Float64,Float64}
    traj::Dict{Int,Int}        # update_sol! copies still in the device trajectory: sol index => slot
    nslots::Int
end

is_rank_adaptive(alg::HIPAlgorithm) = alg.alg isa RankAdaptiveUnconventionalAlgorithm
capacity(alg::HIPAlgorithm, r, n, m) =
    is_rank_adaptive(alg) ? min(alg.alg.alg_params.r_max, LRI_MAX_RANK, n, m) : r

function new_cache(prob, alg::HIPAlgorithm, u)
    n, r = size(u.U)
    m = size(u isa TwoFactorRepresentation ? u.Z : u.V, 1)
    rcap = capacity(alg, r, n, m)
    r <= rcap || throw(ArgumentError("initial rank $r exceeds the capacity $rcap"))
    ctx = HIPContext(n, m, rcap; device = alg.device)
    set_precision!(ctx, alg.precision)
    set_state!(ctx, u)
    t0, tf = prob.tspan
    return HIPCache(ctx, nothing, false, 0, 0, r, u isa TwoFactorRepresentation, 0, 0.0,
                    (Float64(t0), Float64(tf)), Dict{Int,Int}(), 0)
end

# alg_cache(::MatrixDataProblem, ...) (projector_splitting.jl:87-105 and its counterparts)
function alg_cache(prob::MatrixDataProblem, alg::HIPAlgorithm, u, dt; t0 = prob.tspan[1])
    c = new_cache(prob, alg, u)
    c.y = prob.y
    if prob.y isa AbstractArray
        c.array_data = true
        reserve_snapshots!(c.ctx, length(prob.y), alg.snapshot_dtype)
        for (k, A) in enumerate(prob.y)
            upload_snapshot!(c.ctx, k - 1, A, alg.snapshot_dtype)
        end
    else
        reserve_snapshots!(c.ctx, 3, alg.snapshot_dtype)  # y(t), y(t + dt/2), y(t + dt)
        upload_snapshot!(c.ctx, 0, prob.y(t0), alg.snapshot_dtype)
    end
    return c
end

# alg_cache(::MatrixDEProblem, ...) (projector_splitting.jl:43-85, unconventional.jl:39-84,
# rank_adaptive_unconventional.jl:47-92): the operator descriptor goes to the device
function alg_cache(prob::MatrixDEProblem, alg::HIPAlgorithm, u, dt; t0 = prob.tspan[1])
    F = prob.f
    F isa Union{LyapunovOperator,QuadraticOperator} ||
        throw(ArgumentError("a closure cannot run on the device: give f as a LyapunovOperator or QuadraticOperator"))
    c = new_cache(prob, alg, u)
    set_operator!(c.ctx, F)
    use_eigenbasis!(c.ctx, F)
    if alg.substeps !== nothing
        c.substeps = alg.substeps
    elseif F isa QuadraticOperator
        c.quad_ymax = 2 * maximum(norm.(eachrow(u.U * u.S))) * maximum(norm.(eachrow(u.V)))
        c.substeps = rk4_substeps(spectral_bound(F, c.quad_ymax), dt)
    else
        c.substeps = rk4_substeps(spectral_bound(F), dt)
    end
    return c
end

function reserve_trajectory!(c::HIPCache, slots::Integer)
    check(c.ctx.ptr, ccall((:lri_trajectory_reserve, libhip), Cint, (Ptr{Cvoid}, Int64), c.ctx.ptr, slots))
    c.nslots = slots
end

# init(prob, alg, dt) (projector_splitting.jl:107-115)
function init(prob::Union{MatrixDataProblem,MatrixDEProblem}, alg::HIPAlgorithm, dt)
    t0, tf = prob.tspan
    @assert tf > t0 "Integration in reverse time direction is not supported"
    u = deepcopy(prob.u0)
    sol = init_sol(dt, t0, tf, prob.u0)
    cache = alg_cache(prob, alg, u, dt; t0 = t0)
    sol.Y[1] = deepcopy(prob.u0)
    reserve_trajectory!(cache, length(sol.Y) + 2)  # the accumulated t can add a step past init_sol's slots
    return DLRIntegrator(u, t0, dt, sol, alg, cache, typeof(prob), 0)
end

# ---------------------------------------------------------------------------- steps

# the snapshot slots of dy = y(t + dt) - yprev (update_data!, data_integrator.jl:22-28)
function next_snapshots!(c::HIPCache, alg::HIPAlgorithm, t, dt)
    if c.array_data
        kp, kn = c.kprev, t + dt - 1
        0 <= kn < length(c.y) || throw(BoundsError(c.y, kn + 1))
        c.kprev = kn
        return kp, kn
    end
    kp = c.slot
    kn = first(s for s in 0:2 if s != kp)
    upload_snapshot!(c.ctx, kn, c.y(t + dt), alg.snapshot_dtype)
    c.slot = kn
    return kp, kn
end

ksl_order(::ProjectorSplitting{PrimalLieTrotter}) = LRI_PRIMAL
ksl_order(::ProjectorSplitting{DualLieTrotter}) = LRI_DUAL

function rank_changed!(c::HIPCache, r_new)
    if r_new != c.rank
        println("rank adjusted: new rank = $r_new")  # rank_adaptive_unconventional.jl:230
        c.rank = r_new
    end
end

data_step!(c, alg, inner::ProjectorSplitting{<:Union{PrimalLieTrotter,DualLieTrotter}}, t, dt) =
    (k = next_snapshots!(c, alg, t, dt);
     check(c.ctx.ptr, ccall((:lri_step_ksl, libhip), Cint, (Ptr{Cvoid}, Cint, Int64, Int64),
                            c.ctx.ptr, ksl_order(inner), k[1], k[2])))

function data_step!(c, alg, ::ProjectorSplitting{Strang}, t, dt)
    # update_data!(x, y::AbstractArray, t::Int, dt::Int) has no method for dt/2 (data_integrator.jl:26)
    c.array_data && throw(MethodError(LowRankIntegrators.update_data!, (nothing, c.y, t, dt / 2)))
    kp = c.slot
    km, kn = (s for s in 0:2 if s != kp)
    upload_snapshot!(c.ctx, km, c.y(t + dt / 2), alg.snapshot_dtype)
    upload_snapshot!(c.ctx, kn, c.y(t + dt / 2 + dt / 2), alg.snapshot_dtype)
    check(c.ctx.ptr, ccall((:lri_step_ksl_strang, libhip), Cint, (Ptr{Cvoid}, Int64, Int64, Int64),
                           c.ctx.ptr, kp, km, kn))
    c.slot = kn
end

data_step!(c, alg, ::UnconventionalAlgorithm, t, dt) =
    (k = next_snapshots!(c, alg, t, dt);
     check(c.ctx.ptr, ccall((:lri_step_bug, libhip), Cint, (Ptr{Cvoid}, Int64, Int64), c.ctx.ptr, k[1], k[2])))

function data_step!(c, alg, inner::RankAdaptiveUnconventionalAlgorithm, t, dt)
    kp, kn = next_snapshots!(c, alg, t, dt)
    r_new = Ref{Int64}(0)
    check(c.ctx.ptr, ccall((:lri_step_bug_ra, libhip), Cint, (Ptr{Cvoid}, Int64, Int64, Cdouble, Int64, Ref{Int64}),
                           c.ctx.ptr, kp, kn, inner.alg_params.tol, inner.alg_params.r_max, r_new))
    rank_changed!(c, Int(r_new[]))
end

function data_step!(c, alg, ::GreedyIntegrator, t, dt)
    # greedy_step! reads only X = y(t + dt) (greedy_integrator.jl:84-104)
    _, kn = next_snapshots!(c, alg, t, dt)
    if c.twofactor
        check(c.ctx.ptr, ccall((:lri_step_greedy_twofactor, libhip), Cint, (Ptr{Cvoid}, Int64), c.ctx.ptr, kn))
    else
        check(c.ctx.ptr, ccall((:lri_step_greedy, libhip), Cint, (Ptr{Cvoid}, Int64), c.ctx.ptr, kn))
    end
end

de_step!(c, inner::ProjectorSplitting{PrimalLieTrotter}, dt) =
    check(c.ctx.ptr, ccall((:lri_step_de_ksl, libhip), Cint, (Ptr{Cvoid}, Cint, Cdouble, Int64),
                           c.ctx.ptr, LRI_PRIMAL, dt, c.substeps))
de_step!(c, inner::ProjectorSplitting{DualLieTrotter}, dt) =
    check(c.ctx.ptr, ccall((:lri_step_de_ksl, libhip), Cint, (Ptr{Cvoid}, Cint, Cdouble, Int64),
                           c.ctx.ptr, LRI_DUAL, dt, c.substeps))
de_step!(c, inner::ProjectorSplitting{Strang}, dt) =
    check(c.ctx.ptr, ccall((:lri_step_de_ksl, libhip), Cint, (Ptr{Cvoid}, Cint, Cdouble, Int64),
                           c.ctx.ptr, LRI_STRANG, dt, c.substeps))
de_step!(c, ::UnconventionalAlgorithm, dt) =
    check(c.ctx.ptr, ccall((:lri_step_de_bug, libhip), Cint, (Ptr{Cvoid}, Cdouble, Int64), c.ctx.ptr, dt, c.substeps))
function de_step!(c, inner::RankAdaptiveUnconventionalAlgorithm, dt)
    r_new = Ref{Int64}(0)
    check(c.ctx.ptr, ccall((:lri_step_de_bug_ra, libhip), Cint, (Ptr{Cvoid}, Cdouble, Int64, Cdouble, Int64, Ref{Int64}),
                           c.ctx.ptr, dt, c.substeps, inner.alg_params.tol, inner.alg_params.r_max, r_new))
    rank_changed!(c, Int(r_new[]))
end

"""
    step!(integrator::DLRIntegrator, alg::HIPAlgorithm, dt)

One step on the device (asynchronous except for the rank read-back of the
rank-adaptive step), then `t += dt`, `iter += 1` as the reference's `step!`
methods (projector_splitting.jl:191-211, unconventional.jl:159-164,
rank_adaptive_unconventional.jl:171-180, greedy_integrator.jl:106-111).
`integrator.u` is brought up to date by `update_sol!` at the end of a `solve`
(or by [`sync_state!`](@ref)).
"""
function step!(integrator::DLRIntegrator, alg::HIPAlgorithm, dt)
    c = integrator.cache
    if integrator.probType <: MatrixDEProblem
        de_step!(c, alg.alg, dt)
    else
        data_step!(c, alg, alg.alg, integrator.t, dt)
    end
    integrator.t += dt
    integrator.iter += 1
    return nothing
end

"Copy the device state into `integrator.u` (synchronising)."
function sync_state!(integrator::DLRIntegrator)
    c = integrator.cache
    integrator.u = c.twofactor ? get_state_twofactor(c.ctx) : get_state(c.ctx)
    return integrator.u
end

function trajectory_get(c::HIPCache, slot::Integer)
    r = Ref{Int64}(0)
    ctx = c.ctx
    check(ctx.ptr, ccall((:lri_trajectory_get, libhip), Cint,
        (Ptr{Cvoid}, Int64, Ref{Int64}, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64),
        ctx.ptr, slot, r, C_NULL, ctx.n, C_NULL, 1, C_NULL, ctx.m))
    rr = Int(r[])
    U, S, V = zeros(ctx.n, rr), zeros(rr, rr), zeros(ctx.m, rr)
    check(ctx.ptr, ccall((:lri_trajectory_get, libhip), Cint,
        (Ptr{Cvoid}, Int64, Ref{Int64}, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64),
        ctx.ptr, slot, r, U, ctx.n, S, rr, V, ctx.m))
    return c.twofactor ? TwoFactorRepresentation(U, V) : SVDLikeRepresentation(U, S, V)
end

"""
    trajectory_get_range(c::HIPCache, slot0, count)

The states of slots `slot0 .. slot0+count-1` from one copy (`lri_trajectory_get_range`):
U and V are views into one host buffer in the slots' device layout
(`n*rcap + rcap*rcap + m*rcap` doubles per slot), S a copy of its leading block.
"""
function trajectory_get_range(c::HIPCache, slot0::Integer, count::Integer)
    ctx = c.ctx
    n, m, rc = ctx.n, ctx.m, ctx.rcap
    stride = n * rc + rc * rc + m * rc
    buf = Vector{Float64}(undef, max(count, 1) * stride)
    ranks = zeros(Int64, max(count, 1))
    check(ctx.ptr, ccall((:lri_trajectory_get_range, libhip), Cint,
        (Ptr{Cvoid}, Int64, Int64, Ptr{Int64}, Ptr{Float64}), ctx.ptr, slot0, count, ranks, buf))
    out = Any[]
    for i in 1:count
        o, r = (i - 1) * stride, Int(ranks[i])
        U = reshape(view(buf, o+1:o+n*r), n, r)
        S = reshape(buf[o+n*rc+1:o+n*rc+rc*rc], rc, rc)[1:r, 1:r]
        V = reshape(view(buf, o+n*rc+rc*rc+1:o+n*rc+rc*rc+m*r), m, r)
        push!(out, c.twofactor ? TwoFactorRepresentation(U, V) : SVDLikeRepresentation(U, S, V))
    end
    return out
end

"""
    update_sol!(integrator::DLRIntegrator{<:Any,<:Any,<:HIPAlgorithm})

`update_sol!` (primitives.jl:82-90) with the deep copy queued on the device
(`lri_trajectory_save`); after the step on which `solve`'s loop ends
(`(tf - t)/T <= 1e-8`, primitives.jl:71) the whole trajectory is copied into
`sol.Y` and `integrator.u` is brought up to date.
"""
function update_sol!(integrator::DLRIntegrator{<:Any,<:Any,<:HIPAlgorithm})
    c = integrator.cache
    k = integrator.iter
    sol = integrator.sol
    if k <= length(sol.Y) - 1
        sol.t[k+1] = integrator.t
    else
        push!(sol.Y, integrator.u)  # placeholder, replaced from the trajectory
        push!(sol.t, integrator.t)
    end
    if k < c.nslots
        check(c.ctx.ptr, ccall((:lri_trajectory_save, libhip), Cint, (Ptr{Cvoid}, Int64), c.ctx.ptr, k))
        c.traj[k+1] = k
    else
        sol.Y[k+1] = c.twofactor ? get_state_twofactor(c.ctx) : get_state(c.ctx)
    end
    t0, tf = c.tspan
    if !((tf - integrator.t) / (tf - t0) > 1e-8)
        items = sort!(collect(c.traj))
        if c.ctx.resident  # slots saved in the eigenbasis are transformed back one at a time
            for (idx, slot) in items
                sol.Y[idx] = trajectory_get(c, slot)
            end
        else  # runs of consecutive slots, one copy each
            i = 1
            while i <= length(items)
                j = i
                while j < length(items) && items[j+1][2] == items[j][2] + 1
                    j += 1
                end
                for (t, u) in zip(i:j, trajectory_get_range(c, items[i][2], j - i + 1))
                    sol.Y[items[t][1]] = u
                end
                i = j + 1
            end
        end
        empty!(c.traj)
        sync_state!(integrator)
    end
    return nothing
end

# ---------------------------------------------------------------------------- reverse steps (rrule)

"""
    final_state(prob::MatrixDataProblem, alg::HIPAlgorithm, dt = 1)

`LowRankIntegrators.solve(prob, alg, dt).Y[end]`, differentiable with respect
to `prob.u0` through the `rrule` below: the engine's hand-written reverse
steps (`lri_step_bug_adjoint`, `lri_step_ksl_adjoint`, `lri_step_ksl_dual_adjoint`,
`lri_step_bug_ra_adjoint`)
with checkpointed recomputation.  The reference has no adjoint (Julia AD would
differentiate `unconventional_step!` and its counterparts).
"""
final_state(prob::MatrixDataProblem, alg::HIPAlgorithm, dt = 1) = LowRankIntegrators.solve(prob, alg, dt).Y[end]

function forward_step!(ctx::HIPContext, inner, kp, kn)
    if inner isa UnconventionalAlgorithm
        check(ctx.ptr, ccall((:lri_step_bug, libhip), Cint, (Ptr{Cvoid}, Int64, Int64), ctx.ptr, kp, kn))
    elseif inner isa ProjectorSplitting{PrimalLieTrotter}
        check(ctx.ptr, ccall((:lri_step_ksl, libhip), Cint, (Ptr{Cvoid}, Cint, Int64, Int64), ctx.ptr, LRI_PRIMAL, kp, kn))
    elseif inner isa ProjectorSplitting{DualLieTrotter}
        check(ctx.ptr, ccall((:lri_step_ksl, libhip), Cint, (Ptr{Cvoid}, Cint, Int64, Int64), ctx.ptr, LRI_DUAL, kp, kn))
    elseif inner isa RankAdaptiveUnconventionalAlgorithm
        r = Ref{Int64}(0)
        check(ctx.ptr, ccall((:lri_step_bug_ra, libhip), Cint, (Ptr{Cvoid}, Int64, Int64, Cdouble, Int64, Ref{Int64}),
                             ctx.ptr, kp, kn, inner.alg_params.tol, inner.alg_params.r_max, r))
    else
        throw(ArgumentError("the reverse step is built for UnconventionalAlgorithm, " *
                            "ProjectorSplitting(PrimalLieTrotter()), ProjectorSplitting(DualLieTrotter()) " *
                            "and RankAdaptiveUnconventionalAlgorithm"))
    end
end

# one reverse step: (Ub, Sb, Vb) of the step's result -> those of its start (the state set in ctx)
function reverse_step!(ctx::HIPContext, inner, kp, kn, Ub, Sb, Vb, r_new)
    if inner isa RankAdaptiveUnconventionalAlgorithm
        r = rank_of(ctx)
        w = max(r, r_new)
        Ubw, Sbw, Vbw = zeros(ctx.n, w), zeros(w, w), zeros(ctx.m, w)
        Ubw[:, 1:r_new] .= Ub
        Sbw[1:r_new, 1:r_new] .= Sb
        Vbw[:, 1:r_new] .= Vb
        GC.@preserve Ubw Sbw Vbw check(ctx.ptr, ccall((:lri_step_bug_ra_adjoint, libhip), Cint,
            (Ptr{Cvoid}, Int64, Int64, Cdouble, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64,
             Ptr{Float64}, Int64),
            ctx.ptr, kp, kn, inner.alg_params.tol, inner.alg_params.r_max, r_new, Ubw, ctx.n, Sbw, w, Vbw, ctx.m))
        return Ubw[:, 1:r], Sbw[1:r, 1:r], Vbw[:, 1:r]
    end
    Ub, Sb, Vb = Matrix{Float64}(Ub), Matrix{Float64}(Sb), Matrix{Float64}(Vb)
    if inner isa UnconventionalAlgorithm
        GC.@preserve Ub Sb Vb check(ctx.ptr, ccall((:lri_step_bug_adjoint, libhip), Cint,
            (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64),
            ctx.ptr, kp, kn, Ub, size(Ub, 1), Sb, size(Sb, 1), Vb, size(Vb, 1)))
    elseif inner isa ProjectorSplitting{DualLieTrotter}
        GC.@preserve Ub Sb Vb check(ctx.ptr, ccall((:lri_step_ksl_dual_adjoint, libhip), Cint,
            (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64),
            ctx.ptr, kp, kn, Ub, size(Ub, 1), Sb, size(Sb, 1), Vb, size(Vb, 1)))
    else
        GC.@preserve Ub Sb Vb check(ctx.ptr, ccall((:lri_step_ksl_adjoint, libhip), Cint,
            (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64),
            ctx.ptr, kp, kn, Ub, size(Ub, 1), Sb, size(Sb, 1), Vb, size(Vb, 1)))
    end
    return Ub, Sb, Vb
end

function ChainRulesCore.rrule(::typeof(final_state), prob::MatrixDataProblem, alg::HIPAlgorithm, dt = 1;
                              checkpoint::Integer = 20)
    prob.y isa AbstractArray || throw(ArgumentError("the reverse steps take array-form data"))
    inner = alg.alg
    u0 = prob.u0
    n, r = size(u0.U)
    m = size(u0.V, 1)
    ctx = HIPContext(n, m, capacity(alg, r, n, m); device = alg.device)
    set_precision!(ctx, alg.precision)
    set_state!(ctx, u0)
    reserve_snapshots!(ctx, length(prob.y), alg.snapshot_dtype)
    for (k, A) in enumerate(prob.y)
        upload_snapshot!(ctx, k - 1, A, alg.snapshot_dtype)
    end
    t0, tf = prob.tspan
    T = tf - t0
    kprev(s) = s == 0 ? 0 : t0 - 1 + s  # yprev starts at y[1] whatever t0 is
    knext(s) = t0 + s
    cps = Dict(0 => get_factors(ctx))
    ranks = Int[rank_of(ctx)]
    for s in 0:T-1
        forward_step!(ctx, inner, kprev(s), knext(s))
        push!(ranks, rank_of(ctx))
        (s + 1) % checkpoint == 0 && s + 1 < T && (cps[s+1] = get_factors(ctx))
    end
    yT = get_state(ctx)
    function final_state_pullback(ȳ)
        Δ = unthunk(ȳ)
        Ub = Matrix{Float64}(Δ.U isa AbstractZero ? zero(yT.U) : Δ.U)
        Sb = Matrix{Float64}(Δ.S isa AbstractZero ? zero(yT.S) : Δ.S)
        Vb = Matrix{Float64}(Δ.V isa AbstractZero ? zero(yT.V) : Δ.V)
        for c0 in sort!(collect(keys(cps)); rev = true)
            c1 = min(c0 + checkpoint, T)
            states = [cps[c0]]
            set_state!(ctx, cps[c0]...)
            for s in c0:c1-2
                forward_step!(ctx, inner, kprev(s), knext(s))
                push!(states, get_factors(ctx))
            end
            for s in c1-1:-1:c0
                set_state!(ctx, states[s-c0+1]...)
                Ub, Sb, Vb = reverse_step!(ctx, inner, kprev(s), knext(s), Ub, Sb, Vb, ranks[s+2])
            end
        end
        ū0 = Tangent{typeof(u0)}(; U = Ub, S = Sb, V = Vb)
        return NoTangent(), Tangent{typeof(prob)}(; u0 = ū0), NoTangent(), NoTangent()
    end
    return yT, final_state_pullback
end

"""
    solve_gradient(prob::MatrixDEProblem, alg::HIPAlgorithm, dt, loss_grad; checkpoint = 20)

Gradient of `loss_grad(U, S, V) -> (value, Ub, Sb, Vb)` at the final state of
`solve(prob, alg, dt)` for a matrix ODE with a `QuadraticOperator` (D = 0 for a
linear one) and `UnconventionalAlgorithm()`, `RankAdaptiveUnconventionalAlgorithm`
(the ranks of the forward solve held fixed) or `ProjectorSplitting(PrimalLieTrotter())`:
returns `(value, (Ub0, Sb0, Vb0), (abar, bbar))`, abar and bbar the derivatives
along the scales of the operator's A and B (A = ν Laplacian: dL/dν = abar / ν),
through `lri_step_de_bug_adjoint` / `lri_step_de_bug_ra_adjoint` /
`lri_step_de_ksl_adjoint` with checkpointed recomputation.
"""
function solve_gradient(prob::MatrixDEProblem, alg::HIPAlgorithm, dt, loss_grad; checkpoint::Integer = 20)
    inner = alg.alg
    ksl = inner isa ProjectorSplitting{PrimalLieTrotter}
    ra = inner isa RankAdaptiveUnconventionalAlgorithm
    (ksl || ra || inner isa UnconventionalAlgorithm) ||
        throw(ArgumentError("the matrix-ODE adjoint is built for UnconventionalAlgorithm, " *
                            "RankAdaptiveUnconventionalAlgorithm and primal KSL"))
    prob.f isa QuadraticOperator || throw(ArgumentError("the matrix-ODE adjoint takes a QuadraticOperator"))
    c = alg_cache(prob, alg, deepcopy(prob.u0), dt)
    ctx, N = c.ctx, c.substeps
    t0, tf = prob.tspan
    steps, t = 0, t0
    while (tf - t) / (tf - t0) > 1e-8
        t += dt
        steps += 1
    end
    function fwd!()
        if ksl
            check(ctx.ptr, ccall((:lri_step_de_ksl, libhip), Cint, (Ptr{Cvoid}, Cint, Cdouble, Int64),
                                 ctx.ptr, LRI_PRIMAL, dt, N))
        elseif ra
            rn = Ref{Int64}(0)
            check(ctx.ptr, ccall((:lri_step_de_bug_ra, libhip), Cint,
                                 (Ptr{Cvoid}, Cdouble, Int64, Cdouble, Int64, Ref{Int64}),
                                 ctx.ptr, dt, N, inner.alg_params.tol, inner.alg_params.r_max, rn))
        else
            check(ctx.ptr, ccall((:lri_step_de_bug, libhip), Cint, (Ptr{Cvoid}, Cdouble, Int64), ctx.ptr, dt, N))
        end
        return nothing
    end
    cps = Dict(0 => get_factors(ctx))
    for s in 0:steps-1
        fwd!()
        (s + 1) % checkpoint == 0 && s + 1 < steps && (cps[s+1] = get_factors(ctx))
    end
    value, Ub, Sb, Vb = loss_grad(get_factors(ctx)...)
    Ub, Sb, Vb = Matrix{Float64}(Ub), Matrix{Float64}(Sb), Matrix{Float64}(Vb)
    scale = zeros(2)
    for c0 in sort!(collect(keys(cps)); rev = true)
        c1 = min(c0 + checkpoint, steps)
        states = [cps[c0]]
        set_state!(ctx, cps[c0]...)
        for s in c0:c1-2
            fwd!()
            push!(states, get_factors(ctx))
        end
        for s in c1-1:-1:c0
            set_state!(ctx, states[s-c0+1]...)
            if ra  # the new rank r_new is the cotangents' width; the arrays hold max(r, r_new) columns
                r, rn = rank_of(ctx), size(Ub, 2)
                w = max(r, rn)
                Ubw, Sbw, Vbw = zeros(ctx.n, w), zeros(w, w), zeros(ctx.m, w)
                Ubw[:, 1:rn] .= Ub
                Sbw[1:rn, 1:rn] .= Sb
                Vbw[:, 1:rn] .= Vb
                GC.@preserve Ubw Sbw Vbw scale check(ctx.ptr, ccall((:lri_step_de_bug_ra_adjoint, libhip), Cint,
                    (Ptr{Cvoid}, Cdouble, Int64, Cdouble, Int64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64,
                     Ptr{Float64}, Int64, Ptr{Float64}),
                    ctx.ptr, dt, N, inner.alg_params.tol, inner.alg_params.r_max, rn, Ubw, ctx.n, Sbw, w, Vbw, ctx.m,
                    scale))
                Ub, Sb, Vb = Ubw[:, 1:r], Sbw[1:r, 1:r], Vbw[:, 1:r]
            elseif ksl
                GC.@preserve Ub Sb Vb scale check(ctx.ptr, ccall((:lri_step_de_ksl_adjoint, libhip), Cint,
                    (Ptr{Cvoid}, Cdouble, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64,
                     Ptr{Float64}), ctx.ptr, dt, N, Ub, size(Ub, 1), Sb, size(Sb, 1), Vb, size(Vb, 1), scale))
            else
                GC.@preserve Ub Sb Vb scale check(ctx.ptr, ccall((:lri_step_de_bug_adjoint, libhip), Cint,
                    (Ptr{Cvoid}, Cdouble, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64,
                     Ptr{Float64}), ctx.ptr, dt, N, Ub, size(Ub, 1), Sb, size(Sb, 1), Vb, size(Vb, 1), scale))
            end
        end
    end
    return value, (Ub, Sb, Vb), (scale[1], scale[2])
end

# ---------------------------------------------------------------------------- autotuning

"""
    autotune!(c::HIPContext, kp, kn; reps = 10)

`lri_autotune_passes`: time every candidate reduction chunk of the two
streaming passes over A[kn] - A[kp] (0-based slots) at the context's rank and
precision (three interleaved rounds, the median per candidate) and keep the
fastest for this context, the default unless beaten by 1 %.  Returns, per pass,
the chosen chunk, its split count, and its and the default chunk's median time (us).
"""
function autotune!(c::HIPContext, kp::Integer, kn::Integer; reps::Integer = 10)
    rep = zeros(8)
    check(c.ptr, ccall((:lri_autotune_passes, libhip), Cint, (Ptr{Cvoid}, Int64, Int64, Int64, Ptr{Float64}),
                       c.ptr, kp, kn, reps, rep))
    pass(o) = (chunk = Int(rep[o+1]), splits = Int(rep[o+2]), us = rep[o+3], default_us = rep[o+4])
    return (pass_n = pass(0), pass_t = pass(4))
end

# ---------------------------------------------------------------------------- batches

function batch_check(b::Ptr{Cvoid}, st::Cint)
    st == LRI_OK && return nothing
    msg = unsafe_string(ccall((:lri_batch_last_error, libhip), Cstring, (Ptr{Cvoid},), b))
    st == LRI_EINVAL && throw(ArgumentError(msg))
    throw(LRIError(st, msg))
end

"""
    solve_batch(probs, alg::HIPAlgorithm, dt) -> Vector{DLRSolution}

Many independent `MatrixDEProblem`s of uneven sizes and ranks with one `tspan`
stepped together through an `lri_batch`, one context per problem (the
reference solves them one `solve` call at a time, primitives.jl:68-80).
`alg.alg` is `UnconventionalAlgorithm()` or `RankAdaptiveUnconventionalAlgorithm`.
Members whose operator has constant coefficients keep their state in its
eigenbasis; when all do, each stage of a rank-adaptive batch step is one launch
over all members.  Each solution holds the initial and the final state.
"""
function solve_batch(probs::AbstractVector, alg::HIPAlgorithm, dt)
    isempty(probs) && return DLRSolution[]
    inner = alg.alg
    ra = inner isa RankAdaptiveUnconventionalAlgorithm
    (ra || inner isa UnconventionalAlgorithm) ||
        throw(ArgumentError("solve_batch runs UnconventionalAlgorithm and RankAdaptiveUnconventionalAlgorithm"))
    t0, tf = probs[1].tspan
    all(p -> p.tspan == (t0, tf), probs) || throw(ArgumentError("all problems of a batch share tspan"))
    @assert tf > t0 "Integration in reverse time direction is not supported"
    count = length(probs)
    ns = Int64[size(p.u0.U, 1) for p in probs]
    ms = Int64[size(p.u0.V, 1) for p in probs]
    rcaps = Int64[capacity(alg, size(probs[i].u0.U, 2), ns[i], ms[i]) for i in 1:count]
    out = Ref{Ptr{Cvoid}}(C_NULL)
    batch_check(C_NULL, ccall((:lri_batch_create, libhip), Cint,
                              (Ref{Ptr{Cvoid}}, Cint, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}),
                              out, alg.device, count, ns, ms, rcaps))
    b = out[]
    try
        members = MemberContext[]
        substeps = zeros(Int64, count)
        for (i, p) in enumerate(probs)
            cp = Ref{Ptr{Cvoid}}(C_NULL)
            batch_check(b, ccall((:lri_batch_context, libhip), Cint, (Ptr{Cvoid}, Int64, Ref{Ptr{Cvoid}}),
                                 b, i - 1, cp))
            c = MemberContext(cp[], ns[i], ms[i], rcaps[i])
            push!(members, c)
            F = p.f
            F isa Union{LyapunovOperator,QuadraticOperator} ||
                throw(ArgumentError("give f as a LyapunovOperator or QuadraticOperator"))
            set_state!(c, p.u0)
            set_operator!(c, F)
            use_eigenbasis!(c, F)
            substeps[i] = alg.substeps !== nothing ? alg.substeps :
                          F isa QuadraticOperator ?
                          rk4_substeps(spectral_bound(F, 2 * maximum(norm.(eachrow(p.u0.U * p.u0.S))) *
                                                            maximum(norm.(eachrow(p.u0.V)))), dt) :
                          rk4_substeps(spectral_bound(F), dt)
        end
        r_new = zeros(Int64, count)
        t = t0
        while (tf - t) / (tf - t0) > 1e-8
            if ra
                batch_check(b, ccall((:lri_batch_step_de_bug_ra, libhip), Cint,
                                     (Ptr{Cvoid}, Cdouble, Ptr{Int64}, Cdouble, Int64, Ptr{Int64}),
                                     b, dt, substeps, inner.alg_params.tol, inner.alg_params.r_max, r_new))
            else
                batch_check(b, ccall((:lri_batch_step_de_bug, libhip), Cint, (Ptr{Cvoid}, Cdouble, Ptr{Int64}),
                                     b, dt, substeps))
            end
            t += dt
        end
        batch_check(b, ccall((:lri_batch_synchronize, libhip), Cint, (Ptr{Cvoid},), b))
        sols = DLRSolution[]
        for (i, p) in enumerate(probs)
            push!(sols, DLRSolution([deepcopy(p.u0), get_state(members[i])], [Float64(t0), Float64(t)]))
        end
        return sols
    finally
        ccall((:lri_batch_destroy, libhip), Cint, (Ptr{Cvoid},), b)
    end
end

end # module
