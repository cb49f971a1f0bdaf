"""
    PNJLHip

Julia binding of libpnjl_hip.so (include/pnjl_hip.h) for the reference's
`PNJL.solve` / `polarization_aniso` call sites: a thin `ccall` layer with the
reference's argument meaning, so a driver switches backends with one keyword
(`backend=:hip`).  No Julia toolchain exists in the build image; this module
is the binding a maintainer adds to the reference and is not executed here
(the Python mirror in julia_relaxtime_amd/ is the tested host).

Replaces:
  PNJL.solve(::FixedMu, T_fm, mu_fm; ...)          src/pnjl/solver/ImplicitSolver.jl:211-328
  polarization_aniso(channel, k0, k_norm, ...)      src/relaxtime/PolarizationAniso.jl:22-24
  ThermoDerivatives (mass / thermo / bulk-viscosity derivatives)  ThermoDerivatives.jl:132-466
  solve_with_derivatives(T_fm, mu_fm; order)       src/pnjl/solver/ImplicitSolver.jl:660-701
  calculate_propagator_by_channel(process, channel, ...)  src/relaxtime/TotalPropagator.jl:894-963
  solve_meson_mass(meson, quark_params, thermo_params; ...)  src/relaxtime/MesonMass.jl:241-276
and adds the learned seed (no counterpart in the reference): Surrogate, learned_seed, SurrogateSeed.
"""
module PNJLHip

using Random

const LIB = get(ENV, "PNJL_HIP_LIB", joinpath(@__DIR__, "..", "libpnjl_hip.so"))

# policies / flags (include/pnjl_hip.h)
const SEED_DEFAULT, SEED_HADRON, SEED_QUARK, SEED_MULTI, SEED_EXPLICIT = 0, 1, 2, 3, 4
const SEED_CONTINUITY, SEED_PHASE_AWARE, SEED_SURROGATE, SEED_MULTIROOT3 = 5, 6, 7, 8
const SEED_BRANCH_HADRON, SEED_BRANCH_QUARK = 9, 10      # DualBranchScan branches as chains
# per-point flags (PNJL_PT_* of include/pnjl_hip.h)
const PT_TR_USED, PT_MULTISEED, PT_EXCEPTION, PT_NO_ROOT = 0x01, 0x02, 0x04, 0x08
const PT_CHAIN_RESET, PT_BRANCH_END, PT_SKIPPED = 0x80, 0x200, 0x400
const CHANNEL_P, CHANNEL_S = 1, 2

struct SolveOpts
    xtol::Float64
    ftol::Float64
    residual_norm_max::Float64
    phi_tol::Float64
    max_iter::Cint
    trust_region_fallback::Cint
    auto_multiseed_fallback::Cint
    surrogate_max_steps::Cint
    chain_len::Cint
    surrogate_judge::Cint
end

struct SolveOut
    x::Ptr{Float64}; omega::Ptr{Float64}; pressure::Ptr{Float64}; rho_norm::Ptr{Float64}
    entropy::Ptr{Float64}; energy::Ptr{Float64}; masses::Ptr{Float64}; n_q::Ptr{Float64}
    n_qbar::Ptr{Float64}; residual_norm::Ptr{Float64}; iterations::Ptr{Cint}; converged::Ptr{Cint}
    flags::Ptr{Cint}; newton_steps::Ptr{Cint}; passes::Ptr{Cint}
end

struct PolQuery
    channel::Cint; num_s::Cint; width::Cint; reserved::Cint
    k0::Float64; gamma::Float64; k::Float64; m1::Float64; m2::Float64; mu1::Float64; mu2::Float64
    T::Float64; Phi::Float64; Phib::Float64; xi::Float64; A1::Float64; A2::Float64
end

mutable struct Context
    h::Ptr{Cvoid}
end

function check(ctx::Context, rc::Cint, what)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:pnjl_last_error, LIB), Cstring, (Ptr{Cvoid},), ctx.h))
    error("$what failed with status $rc: $msg")
end

"""One context per device and host thread (calls on a context are serialised)."""
function Context(device::Integer=0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:pnjl_ctx_create, LIB), Cint, (Ptr{Cvoid}, Cint, Ref{Ptr{Cvoid}}), C_NULL, device, h)
    rc == 0 || error("pnjl_ctx_create failed with status $rc")
    ctx = Context(h[])
    finalizer(c -> ccall((:pnjl_ctx_destroy, LIB), Cvoid, (Ptr{Cvoid},), c.h), ctx)
    return ctx
end

const DEFAULT_CTX = Ref{Union{Nothing,Context}}(nothing)
default_context() = something(DEFAULT_CTX[], (DEFAULT_CTX[] = Context(0)))

"""
    solve_batch(T_fm, mu_fm; xi, policy, seeds, p_num=64, t_num=8, iterations=1000, ...)

Batched `PNJL.solve(FixedMu(), T, mu; ...)` in one persistent-kernel launch.
Returns a NamedTuple of arrays (SolverResult fields, point-major).
"""
function solve_batch(T_fm::AbstractVector{Float64}, mu_fm::AbstractVector{Float64};
                     xi::AbstractVector{Float64}=zeros(length(T_fm)), policy::Integer=SEED_DEFAULT,
                     seeds::Union{Nothing,Matrix{Float64}}=nothing, p_num::Integer=64, t_num::Integer=8,
                     iterations::Integer=1000, trust_region_fallback::Bool=true,
                     auto_multiseed_fallback::Bool=true, residual_norm_max::Float64=1e-6,
                     xtol::Float64=1e-9, ftol::Float64=1e-9, chain_len::Integer=0,
                     surrogate_max_steps::Integer=8, surrogate_judge::Integer=0,
                     ctx::Context=default_context())
    n = length(T_fm)
    opts = SolveOpts(xtol, ftol, residual_norm_max, 1e-8, iterations, trust_region_fallback,
                     auto_multiseed_fallback, surrogate_max_steps, chain_len, surrogate_judge)
    x = zeros(5, n); om = zeros(n); P = zeros(n); rho = zeros(n); s = zeros(n); e = zeros(n)
    M = zeros(3, n); nq = zeros(3, n); nqb = zeros(3, n); res = zeros(n)
    it = zeros(Cint, n); conv = zeros(Cint, n); fl = zeros(Cint, n); st = zeros(Cint, n); ps = zeros(Cint, n)
    GC.@preserve T_fm mu_fm xi seeds x om P rho s e M nq nqb res it conv fl st ps begin
        out = SolveOut(pointer(x), pointer(om), pointer(P), pointer(rho), pointer(s), pointer(e), pointer(M),
                       pointer(nq), pointer(nqb), pointer(res), pointer(it), pointer(conv), pointer(fl),
                       pointer(st), pointer(ps))
        sp = seeds === nothing ? Ptr{Float64}(C_NULL) : pointer(seeds)   # 5 x n, column-major = point-major
        rc = ccall((:pnjl_solve_fixedmu_batch, LIB), Cint,
                   (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cint, Cint, Cint,
                    Ref{SolveOpts}, Ref{SolveOut}, Cint, Ptr{Cvoid}),
                   ctx.h, n, T_fm, mu_fm, xi, sp, policy, p_num, t_num, opts, out, 0, C_NULL)
        check(ctx, rc, "pnjl_solve_fixedmu_batch")
    end
    return (solution=x, omega=om, pressure=P, rho_norm=rho, entropy=s, energy=e, masses=M, n_q=nq,
            n_qbar=nqb, residual_norm=res, iterations=it, converged=conv .!= 0, flags=fl)
end

# seed constants of the GPU's DefaultSeed / MultiSeed policies (SeedStrategies.jl:56-107, 257-268)
const HADRON_SEED = [-1.84329, -1.84329, -2.22701, 1.0e-5, 4.0e-5]
const QUARK_SEED = [-0.73192, -0.73192, -1.79539, 0.60532, 0.60532]
const WEAK_CHIRAL_CONF_SEED = [-0.5, -0.5, -1.2, 1e-3, 1e-3]
const HT_0P8_SEED = [-0.5, -0.5, -1.2, 0.8, 0.8]
const HT_0P9_SEED = [-0.3, -0.3, -0.9, 0.9, 0.9]
const HT_0P95_SEED = [-0.2, -0.2, -0.7, 0.95, 0.95]
const MULTI_CANDIDATES = ((HADRON_SEED, QUARK_SEED, :hadron), (HADRON_SEED, QUARK_SEED, :quark),
                          (WEAK_CHIRAL_CONF_SEED, WEAK_CHIRAL_CONF_SEED, :hadron),
                          (HT_0P8_SEED, HT_0P8_SEED, :hadron), (HT_0P9_SEED, HT_0P9_SEED, :hadron),
                          (HT_0P95_SEED, HT_0P95_SEED, :hadron))

# MultiSeed() with its default candidates and Omega selector (the GPU's SEED_MULTI)
is_default_multiseed(s) = length(s.candidates) == 6 && nameof(s.selector) === :default_omega_selector &&
    all(nameof(typeof(c)) === :DefaultSeed && c.hadron_seed == h && c.quark_seed == q && c.phase_hint === hint
        for (c, (h, q, hint)) in zip(s.candidates, MULTI_CANDIDATES))

"""
Fields and order of the reference's `SolverResult` (ImplicitSolver.jl:178-193);
`x_state`, `mu_vec` and `masses` are tuples here (the reference's SVectors
convert from them, so `make_result=SolverResult` builds the reference's own type).
"""
struct SolverResult
    mode::Any
    converged::Bool
    solution::Vector{Float64}
    x_state::NTuple{5,Float64}
    mu_vec::NTuple{3,Float64}
    omega::Float64
    pressure::Float64
    rho_norm::Float64
    entropy::Float64
    energy::Float64
    masses::NTuple{3,Float64}
    iterations::Int
    residual_norm::Float64
    xi::Float64
end

# policy and explicit seed of a seed strategy, by the reference's type names
# and fields (SeedStrategies.jl:193-888); strategies with host-side state the
# GPU does not model (phase-aware seeds, custom seed vectors) take their seed
# from `get_seed` (the reference's get_seed(strategy, [T, mu], mode)) as an
# explicit seed
function policy_for(s, T_fm, mu_fm, mode, get_seed)
    s === nothing && return SEED_DEFAULT, nothing
    t = nameof(typeof(s))
    if t === :DefaultSeed && s.hadron_seed == HADRON_SEED && s.quark_seed == QUARK_SEED
        return (s.phase_hint === :hadron ? SEED_HADRON : s.phase_hint === :quark ? SEED_QUARK : SEED_DEFAULT), nothing
    elseif t === :MultiSeed
        is_default_multiseed(s) && return SEED_MULTI, nothing
        throw(ArgumentError("MultiSeed: the HIP backend runs the default candidates and Omega selector only"))
    elseif t === :ContinuitySeed
        s.previous_solution === nothing && return policy_for(s.fallback, T_fm, mu_fm, mode, get_seed)
        return SEED_EXPLICIT, Float64.(s.previous_solution[1:5])
    elseif t === :PhaseAwareContinuitySeed && s.bootstrap_multiseed && s.previous_solution === nothing
        return SEED_MULTI, nothing
    end
    get_seed === nothing && throw(ArgumentError(
        "seed strategy $t: pass get_seed (the reference's get_seed) so its seed can be used as an explicit seed"))
    return SEED_EXPLICIT, Float64.(get_seed(s, [Float64(T_fm), Float64(mu_fm)], mode)[1:5])
end

"""
    solve(mode, T_fm, mu_fm; xi=0.0, seed_strategy=nothing, p_num=64, t_num=8, iterations=1000,
          trust_region_fallback=true, auto_multiseed_fallback=true, residual_norm_max=1e-6,
          xtol=1e-9, ftol=1e-9, get_seed=nothing, make_result=SolverResult) -> SolverResult

`PNJL.solve(::FixedMu, T_fm, μ_fm; ...)` (ImplicitSolver.jl:211-328) on the GPU,
with the reference's keywords and result.  Non-convergence is `converged = false`;
a non-finite residual at the seed throws `DomainError` (the reference's
NLsolve IsFiniteException) and MultiSeed without a converged candidate throws
the reference's "All seeds failed" error (ImplicitSolver.jl:553-556).
The reference's solve switches with one keyword:

    backend === :hip && return PNJLHip.solve(mode, T_fm, μ_fm; make_result=SolverResult,
                                             get_seed=get_seed, kwargs...)
"""
function solve(mode, T_fm::Real, mu_fm::Real; xi::Real=0.0, seed_strategy=nothing, p_num::Integer=64,
               t_num::Integer=8, iterations::Integer=1000, trust_region_fallback::Bool=true,
               auto_multiseed_fallback::Bool=true, residual_norm_max::Real=1e-6, xtol::Real=1e-9,
               ftol::Real=1e-9, nlsolve_method::Symbol=:newton, fallback_method::Symbol=:trust_region,
               physicality_check=nothing, get_seed=nothing, make_result=SolverResult,
               ctx::Context=default_context())
    nameof(typeof(mode)) === :FixedMu || throw(ArgumentError("the HIP backend solves FixedMu() only"))
    nlsolve_method === :newton || throw(ArgumentError("the HIP backend runs nlsolve_method=:newton"))
    fallback_method === :trust_region || throw(ArgumentError("the HIP backend falls back to :trust_region"))
    physicality_check === nothing || nameof(physicality_check) === :_default_is_physical_solution ||
        throw(ArgumentError("the HIP backend applies the default physicality check"))
    seed_strategy isa SurrogateSeed && return solve_learned(mode, [Float64(T_fm)], [Float64(mu_fm)], seed_strategy;
                                                            xi=[Float64(xi)], p_num=p_num, t_num=t_num,
                                                            iterations=iterations, make_result=make_result,
                                                            ctx=ctx, trust_region_fallback=trust_region_fallback,
                                                            auto_multiseed_fallback=auto_multiseed_fallback,
                                                            residual_norm_max=Float64(residual_norm_max),
                                                            xtol=Float64(xtol), ftol=Float64(ftol))[1]
    policy, x0 = policy_for(seed_strategy, T_fm, mu_fm, mode, get_seed)
    r = solve_batch([Float64(T_fm)], [Float64(mu_fm)]; xi=[Float64(xi)], policy=policy,
                    seeds=x0 === nothing ? nothing : reshape(x0, 5, 1), p_num=p_num, t_num=t_num,
                    iterations=iterations, trust_region_fallback=trust_region_fallback,
                    auto_multiseed_fallback=auto_multiseed_fallback, residual_norm_max=Float64(residual_norm_max),
                    xtol=Float64(xtol), ftol=Float64(ftol), ctx=ctx)
    fl = r.flags[1]
    fl & PT_EXCEPTION != 0 && throw(DomainError(T_fm, "non-finite residual at the seed (NLsolve IsFiniteException)"))
    fl & PT_NO_ROOT != 0 && error("All seeds failed to converge to a physical solution")
    x = r.solution[:, 1]
    mu = Float64(mu_fm)
    return make_result(mode, r.converged[1], copy(x), Tuple(x), (mu, mu, mu), r.omega[1], r.pressure[1],
                       r.rho_norm[1], r.entropy[1], r.energy[1], Tuple(r.masses[:, 1]), Int(r.iterations[1]),
                       r.residual_norm[1], Float64(xi))
end

"""
    solve_many(mode, T_fm, mu_fm; xi, seed_strategy=nothing, ..., make_result=SolverResult)

`solve` over arrays of points in one launch (independent points; DefaultSeed
or MultiSeed strategies): a Vector of results.  Scans that carry a
continuation seed from point to point run as chains with `solve_batch`
(policy SEED_CONTINUITY / SEED_PHASE_AWARE, `chain_len`).
"""
function solve_many(mode, T_fm::AbstractVector{Float64}, mu_fm::AbstractVector{Float64};
                    xi::AbstractVector{Float64}=zeros(length(T_fm)), seed_strategy=nothing, p_num::Integer=64,
                    t_num::Integer=8, iterations::Integer=1000, make_result=SolverResult,
                    ctx::Context=default_context(), kwargs...)
    nameof(typeof(mode)) === :FixedMu || throw(ArgumentError("the HIP backend solves FixedMu() only"))
    seed_strategy isa SurrogateSeed && return solve_learned(mode, T_fm, mu_fm, seed_strategy; xi=xi, p_num=p_num,
                                                            t_num=t_num, iterations=iterations,
                                                            make_result=make_result, ctx=ctx, kwargs...)
    policy, x0 = policy_for(seed_strategy, T_fm[1], mu_fm[1], mode, nothing)
    x0 === nothing || throw(ArgumentError("solve_many: point-independent strategies only"))
    r = solve_batch(T_fm, mu_fm; xi=xi, policy=policy, p_num=p_num, t_num=t_num, iterations=iterations,
                    ctx=ctx, kwargs...)
    return [make_result(mode, r.converged[i], r.solution[:, i], Tuple(r.solution[:, i]),
                        (mu_fm[i], mu_fm[i], mu_fm[i]), r.omega[i], r.pressure[i], r.rho_norm[i], r.entropy[i],
                        r.energy[i], Tuple(r.masses[:, i]), Int(r.iterations[i]), r.residual_norm[i], xi[i])
            for i in eachindex(T_fm)]
end

"""
    thermo_derivatives_batch(T_fm, mu_fm, x; xi, p_num=64, t_num=8)

Implicit thermodynamic derivatives at the roots `x` (5 x n, from `solve_batch`):
a 36 x n matrix, columns as documented for `pnjl_thermo_derivatives_batch`
(pressure, energy, rho, rho_norm, entropy, dP_dT, ..., v_n_sq, dmuB_dT_sigma,
dx_dT, dx_dmu, ds_dT, ds_dmu).  Replaces ThermoDerivatives.jl:132-466
(mass_derivatives order 1, thermo_derivatives, bulk_viscosity_coefficients).
"""
function thermo_derivatives_batch(T_fm::AbstractVector{Float64}, mu_fm::AbstractVector{Float64},
                                  x::Matrix{Float64}; xi::AbstractVector{Float64}=zeros(length(T_fm)),
                                  p_num::Integer=64, t_num::Integer=8, ctx::Context=default_context())
    n = length(T_fm)
    out = zeros(36, n)
    GC.@preserve T_fm mu_fm xi x out begin
        rc = ccall((:pnjl_thermo_derivatives_batch, LIB), Cint,
                   (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cint, Cint,
                    Ptr{Float64}, Cint, Ptr{Cvoid}),
                   ctx.h, n, T_fm, mu_fm, xi, x, p_num, t_num, out, 0, C_NULL)
        check(ctx, rc, "pnjl_thermo_derivatives_batch")
    end
    return out
end

"""
    implicit_derivatives2_batch(T_fm, mu_fm, x; xi, p_num=64, t_num=8)

x, its first and second (T, mu) derivatives and those of the masses at the
roots `x` (5 x n): a 48 x n matrix, rows as documented for
`pnjl_implicit_derivatives2_batch` (x, dx_dT, dx_dmu, d2x_dT2, d2x_dTdmu,
d2x_dmu2, masses, dM_dT, dM_dmu, d2M_dT2, d2M_dTdmu, d2M_dmu2).
"""
function implicit_derivatives2_batch(T_fm::AbstractVector{Float64}, mu_fm::AbstractVector{Float64},
                                     x::Matrix{Float64}; xi::AbstractVector{Float64}=zeros(length(T_fm)),
                                     p_num::Integer=64, t_num::Integer=8, ctx::Context=default_context())
    n = length(T_fm)
    out = zeros(48, n)
    GC.@preserve T_fm mu_fm xi x out begin
        rc = ccall((:pnjl_implicit_derivatives2_batch, LIB), Cint,
                   (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cint, Cint,
                    Ptr{Float64}, Cint, Ptr{Cvoid}),
                   ctx.h, n, T_fm, mu_fm, xi, x, p_num, t_num, out, 0, C_NULL)
        check(ctx, rc, "pnjl_implicit_derivatives2_batch")
    end
    return out
end

function _derivs2(T_fm, mu_fm, xi, p_num, t_num, ctx)
    r = solve_batch([Float64(T_fm)], [Float64(mu_fm)]; xi=[Float64(xi)], p_num=p_num, t_num=t_num, ctx=ctx)
    return implicit_derivatives2_batch([Float64(T_fm)], [Float64(mu_fm)], r.solution; xi=[Float64(xi)],
                                       p_num=p_num, t_num=t_num, ctx=ctx)[:, 1]
end

"""
    solve_with_derivatives(T_fm, mu_fm; order=1, xi=0.0, p_num=64, t_num=8)

Same arguments and NamedTuple as the reference (ImplicitSolver.jl:660-701):
(x, dx_dT, dx_dμ) and, at order 2, d2x_dT2, d2x_dμ2, d2x_dTdμ.
"""
function solve_with_derivatives(T_fm::Real, mu_fm::Real; order::Int=1, xi::Real=0.0, p_num::Int=64,
                                t_num::Int=8, ctx::Context=default_context())
    order in (1, 2) || error("order must be 1 or 2, got $order")
    d = _derivs2(T_fm, mu_fm, xi, p_num, t_num, ctx)
    order == 1 && return (x=d[1:5], dx_dT=d[6:10], dx_dμ=d[11:15])
    return (x=d[1:5], dx_dT=d[6:10], dx_dμ=d[11:15], d2x_dT2=d[16:20], d2x_dμ2=d[26:30], d2x_dTdμ=d[21:25])
end

"""
    mass_derivatives(T_fm, mu_fm; order=1, xi=0.0, p_num=64, t_num=8)

Same arguments and NamedTuple as the reference (ThermoDerivatives.jl:132-179):
(masses, dM_dT, dM_dmu) and, at order 2, d2M_dT2, d2M_dTdmu, d2M_dmu2.
"""
function mass_derivatives(T_fm::Real, mu_fm::Real; order::Int=1, xi::Real=0.0, p_num::Int=64, t_num::Int=8,
                          ctx::Context=default_context())
    order in (1, 2) || error("order must be 1 or 2, got $order")
    d = _derivs2(T_fm, mu_fm, xi, p_num, t_num, ctx)
    order == 1 && return (masses=d[31:33], dM_dT=d[34:36], dM_dmu=d[37:39])
    return (masses=d[31:33], dM_dT=d[34:36], dM_dmu=d[37:39], d2M_dT2=d[40:42], d2M_dTdmu=d[43:45],
            d2M_dmu2=d[46:48])
end

# scattering processes in SCATTERING_MESON_MAP's order (PNJL_PROC_* ids 0..16)
const PROCESSES = (:ud_to_ud, :uu_to_uu, :us_to_us, :ss_to_ss, :ubardbar_to_ubardbar, :ubarubar_to_ubarubar,
                   :ubarsbar_to_ubarsbar, :sbarsbar_to_sbarsbar, :udbar_to_udbar, :dubar_to_dubar,
                   :uubar_to_uubar, :uubar_to_ddbar, :usbar_to_usbar, :subar_to_subar, :uubar_to_ssbar,
                   :ssbar_to_uubar, :ssbar_to_ssbar)
const MANDELSTAM = Dict(:t => 0, :s => 1, :u => 2)

struct PropQuery
    process::Cint; channel::Cint; mesons::Cint; reserved::Cint
    k0::Float64; k::Float64; T::Float64; Phi::Float64; Phib::Float64; xi::Float64
    m_u::Float64; m_s::Float64; mu_u::Float64; mu_s::Float64; A_u::Float64; A_s::Float64
end

"""
    propagator_by_channel_batch(procs, channels, k0s, ks, quark_params, thermo_params; K_coeffs=nothing)

(D_S, D_P) vectors for every (process, channel, k0, k): one call evaluates
every Pi the exchanged mesons need and composes the propagators on the GPU.
quark_params / thermo_params as in the reference (m, μ, A of u and s; T, Φ,
Φbar, ξ).  K_coeffs = nothing derives the couplings from quark_params.A.
Replaces calculate_propagator_by_channel  TotalPropagator.jl:894-963.
"""
function propagator_by_channel_batch(procs::AbstractVector{Symbol}, channels::AbstractVector{Symbol},
                                     k0s::AbstractVector{Float64}, ks::AbstractVector{Float64},
                                     quark_params, thermo_params; K_coeffs::Union{Nothing,Matrix{Float64}}=nothing,
                                     ctx::Context=default_context())
    n = length(procs)
    qp, tp = quark_params, thermo_params
    q = [PropQuery(findfirst(==(procs[i]), PROCESSES) - 1, MANDELSTAM[channels[i]], 0, 0, k0s[i], ks[i],
                   tp.T, tp.Φ, tp.Φbar, tp.ξ, qp.m.u, qp.m.s, qp.μ.u, qp.μ.s, qp.A.u, qp.A.s) for i in 1:n]
    out = zeros(4, n)
    kc = K_coeffs === nothing ? Ptr{Float64}(C_NULL) : pointer(K_coeffs)   # 12 x n, column-major = row-major n x 12
    GC.@preserve q out K_coeffs begin
        rc = ccall((:pnjl_propagator_by_channel_batch, LIB), Cint,
                   (Ptr{Cvoid}, Int64, Ptr{PropQuery}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Cvoid}),
                   ctx.h, n, q, kc, out, 0, C_NULL)
        check(ctx, rc, "pnjl_propagator_by_channel_batch")
    end
    return complex.(out[1, :], out[2, :]), complex.(out[3, :], out[4, :])
end

"""
    calculate_propagator_by_channel(process, channel, k0, k_norm, quark_params, thermo_params)

Same arguments and (D_S, D_P) result as the reference (K_coeffs from quark_params.A).
"""
function calculate_propagator_by_channel(process::Symbol, channel::Symbol, k0::Float64, k_norm::Float64,
                                         quark_params, thermo_params; ctx::Context=default_context())
    DS, DP = propagator_by_channel_batch([process], [channel], [k0], [k_norm], quark_params, thermo_params; ctx=ctx)
    return DS[1], DP[1]
end

const MASS_MESONS = (:pi, :K, :sigma_pi, :sigma_K, :eta, :eta_prime, :sigma, :sigma_prime)

struct MassQuery
    meson::Cint; max_iter::Cint; reserved1::Cint; reserved2::Cint
    k::Float64; mass0::Float64; gamma0::Float64; ftol::Float64; xtol::Float64
    T::Float64; Phi::Float64; Phib::Float64; xi::Float64
    m_u::Float64; m_s::Float64; mu_u::Float64; mu_s::Float64; A_u::Float64; A_s::Float64
end

"""
    solve_meson_mass(meson, quark_params, thermo_params; k_norm=0.0, initial_mass=nothing,
                     initial_gamma=0.0, ftol=1e-8, xtol=0.0, iterations=1000)

Same arguments and (mass, gamma, converged, residual_norm) as the reference
(quark_params must carry A; K_coeffs from it).  Replaces MesonMass.jl:241-276.
"""
function solve_meson_mass(meson::Symbol, quark_params, thermo_params; k_norm::Float64=0.0,
                          initial_mass::Union{Nothing,Float64}=nothing, initial_gamma::Float64=0.0,
                          ftol::Float64=1e-8, xtol::Float64=0.0, iterations::Integer=1000,
                          ctx::Context=default_context())
    qp, tp = quark_params, thermo_params
    q = Ref(MassQuery(findfirst(==(meson), MASS_MESONS) - 1, iterations, 0, 0, k_norm,
                      initial_mass === nothing ? NaN : initial_mass, initial_gamma, ftol, xtol,
                      tp.T, tp.Φ, tp.Φbar, hasproperty(tp, :ξ) ? tp.ξ : 0.0,
                      qp.m.u, qp.m.s, qp.μ.u, qp.μ.s, qp.A.u, qp.A.s))
    out = zeros(6)
    rc = ccall((:pnjl_meson_mass_batch, LIB), Cint,
               (Ptr{Cvoid}, Int64, Ref{MassQuery}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Cvoid}),
               ctx.h, 1, q, C_NULL, out, 0, C_NULL)
    check(ctx, rc, "pnjl_meson_mass_batch")
    out[6] == 1.0 && throw(ErrorException("solve_meson_mass: non-finite residual at the initial point"))
    return (mass=out[1], gamma=out[2], converged=out[4] == 1.0, residual_norm=out[3])
end

"""
    polarization_aniso(channel, k0, k_norm, m1, m2, mu1, mu2, T, Phi, Phibar, xi, A1, A2, num_s; backend=:hip)

Same signature and result (polarization_real, polarization_imag) as the
reference (PolarizationAniso.jl:22-24; channel :P or :S, ArgumentError
otherwise); computed on the GPU.  The reference's function switches with one
keyword: `backend === :hip && return PNJLHip.polarization_aniso(args...)`.
"""
function polarization_aniso(channel::Symbol, k0, k_norm, m1, m2, mu1, mu2, T, Phi, Phibar, xi,
                            A1_value, A2_value, num_s_quark::Int; backend::Symbol=:hip,
                            ctx::Context=default_context())
    backend === :hip || throw(ArgumentError("PNJLHip.polarization_aniso computes backend=:hip only"))
    ch = channel === :P ? CHANNEL_P : channel === :S ? CHANNEL_S :
         throw(ArgumentError("Unsupported channel: $channel. Use :P or :S."))
    q = Ref(PolQuery(ch, num_s_quark, 0, 0, k0, 0.0, k_norm, m1, m2, mu1, mu2, T, Phi, Phibar, xi,
                     A1_value, A2_value))
    re = Ref(0.0); im = Ref(0.0)
    rc = ccall((:pnjl_polarization_batch, LIB), Cint,
               (Ptr{Cvoid}, Int64, Ref{PolQuery}, Ref{Float64}, Ref{Float64}, Cint, Ptr{Cvoid}),
               ctx.h, 1, q, re, im, 0, C_NULL)
    check(ctx, rc, "pnjl_polarization_batch")
    return re[], im[]
end

# ---------------------------------------------------------------------------
# Learned seed (no counterpart in the reference): the MLP seed model of
# pnjl_surrogate_*, its judge mask, and SurrogateSeed, the seed strategy that
# runs PNJL_SEED_SURROGATE / PNJL_SEED_MULTIROOT3.  The reference seeds stay
# the fallback and the judge (SeedStrategies.jl:204-284).  Same recipe as the
# Python host's surrogate.learned_seed (training box, Adam steps, judge-mask
# probes); the random points come from Julia's RNG, so the trained network
# is not the Python one's bit for bit.
# ---------------------------------------------------------------------------
const JUDGE_REFERENCE, JUDGE_OMEGA = 0, 1
const PT_SURR_FALLBACK, PT_SURR_OMEGA, PT_FIRST_ORDER, PT_SURR_MASKED = 0x10, 0x20, 0x40, 0x100
const DOMAIN_LO = [20.0 / 197.327, 0.0, -0.4]                    # T, mu (fm^-1), xi of the training box
const DOMAIN_HI = [300.0 / 197.327, 450.0 / 197.327, 0.4]

struct LaunchCfg
    lanes_per_point::Cint
    chunk::Cint
    waves_per_cu::Cint
    tail_lanes_per_point::Cint
end
# the set-up solves' fixed geometry (the trained network must not depend on the scans' geometry)
const SETUP_LAUNCH = LaunchCfg(8, 1, 16, 0)

function launch_cfg(ctx::Context=default_context())
    c = Ref(LaunchCfg(0, 0, 0, 0))
    check(ctx, ccall((:pnjl_get_launch_cfg, LIB), Cint, (Ptr{Cvoid}, Ref{LaunchCfg}), ctx.h, c),
          "pnjl_get_launch_cfg")
    return c[]
end

set_launch_cfg!(cfg::LaunchCfg, ctx::Context=default_context()) =
    check(ctx, ccall((:pnjl_set_launch_cfg, LIB), Cint, (Ptr{Cvoid}, Ref{LaunchCfg}), ctx.h, cfg),
          "pnjl_set_launch_cfg")

mutable struct Surrogate
    h::Ptr{Cvoid}
    in_lo::Vector{Float64}
    in_hi::Vector{Float64}
    out_mean::Vector{Float64}
    out_std::Vector{Float64}
    judge_mask::Union{Nothing,Array{UInt8,3}}     # [nxi, nmu, nT]: the C layout (T slowest) in column-major
end

function scheck(s::Surrogate, rc::Cint, what)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:pnjl_surrogate_last_error, LIB), Cstring, (Ptr{Cvoid},), s.h))
    error("$what failed with status $rc: $msg")
end

"""A GPU-resident MLP seed model (3 -> 128 -> 128 -> 128 -> 4, BF16 MFMA, FP32 Adam)."""
function Surrogate(device::Integer=0; init_seed::Integer=4)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:pnjl_surrogate_create, LIB), Cint, (Cint, UInt64, Ref{Ptr{Cvoid}}), device, init_seed, h)
    rc == 0 || error("pnjl_surrogate_create failed with status $rc")
    s = Surrogate(h[], copy(DOMAIN_LO), copy(DOMAIN_HI), zeros(4), ones(4), nothing)
    finalizer(x -> ccall((:pnjl_surrogate_destroy, LIB), Cvoid, (Ptr{Cvoid},), x.h), s)
    set_scaling!(s)
    return s
end

set_scaling!(s::Surrogate) =
    scheck(s, ccall((:pnjl_surrogate_set_scaling, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                    s.h, s.in_lo, s.in_hi, s.out_mean, s.out_std), "pnjl_surrogate_set_scaling")

"""Output normalisation from training roots (5 x n): mean and std of (phi_u, phi_s, Phi, Phibar)."""
function fit_output_scaling!(s::Surrogate, roots::AbstractMatrix{Float64})
    y = roots[[1, 3, 4, 5], :]
    n = size(y, 2)
    s.out_mean = vec(sum(y; dims=2)) ./ n
    s.out_std = max.(sqrt.(vec(sum((y .- s.out_mean) .^ 2; dims=2)) ./ n), 1e-6)
    set_scaling!(s)
end

param_count() = Int(ccall((:pnjl_surrogate_param_count, LIB), Cint, ()))

function params(s::Surrogate)
    p = zeros(Float32, param_count())
    scheck(s, ccall((:pnjl_surrogate_get_params, LIB), Cint, (Ptr{Cvoid}, Ptr{Float32}), s.h, p),
           "pnjl_surrogate_get_params")
    return p
end

function params!(s::Surrogate, p::Vector{Float32})
    length(p) == param_count() || throw(ArgumentError("expected $(param_count()) parameters"))
    scheck(s, ccall((:pnjl_surrogate_set_params, LIB), Cint, (Ptr{Cvoid}, Ptr{Float32}), s.h, p),
           "pnjl_surrogate_set_params")
end

"""Adam on the GPU over (T, mu, xi) -> roots (5 x n); returns the per-step loss."""
function train!(s::Surrogate, T_fm::Vector{Float64}, mu_fm::Vector{Float64}, xi::Vector{Float64},
                roots::Matrix{Float64}; steps::Integer=8000, batch::Integer=8192, lr::Real=2e-3, seed::Integer=4)
    loss = zeros(Float32, max(steps, 1))
    scheck(s, ccall((:pnjl_surrogate_train, LIB), Cint,
                    (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cint, Cint,
                     Float32, UInt64, Ptr{Float32}, Cint, Ptr{Cvoid}),
                    s.h, length(T_fm), T_fm, mu_fm, xi, roots, steps, batch, Float32(lr), seed, loss, 0, C_NULL),
           "pnjl_surrogate_train")
    return loss[1:steps]
end

"""Seeds (5 x n: phi_u, phi_u, phi_s, Phi, Phibar) at (T, mu, xi)."""
function infer(s::Surrogate, T_fm::Vector{Float64}, mu_fm::Vector{Float64}, xi::Vector{Float64})
    seeds = zeros(5, length(T_fm))
    scheck(s, ccall((:pnjl_surrogate_infer, LIB), Cint,
                    (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Cvoid}),
                    s.h, length(T_fm), T_fm, mu_fm, xi, seeds, 0, C_NULL), "pnjl_surrogate_infer")
    return seeds
end

"""Installs the model's judge mask in the context (nothing clears it)."""
function set_judge_mask!(ctx::Context, s::Union{Nothing,Surrogate})
    if s === nothing || s.judge_mask === nothing
        rc = ccall((:pnjl_set_judge_mask, LIB), Cint,
                   (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{UInt8}),
                   ctx.h, 0, 0, 0, C_NULL, C_NULL, C_NULL)
    else
        nxi, nmu, nT = size(s.judge_mask)
        rc = ccall((:pnjl_set_judge_mask, LIB), Cint,
                   (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{UInt8}),
                   ctx.h, nT, nmu, nxi, s.in_lo, s.in_hi, s.judge_mask)
    end
    check(ctx, rc, "pnjl_set_judge_mask")
end

# DualBranchScan's sameness test (DualBranchScan.jl:409-419) and Omega to 1e-10
same_root(a, b, i) = abs(a.solution[1, i] - b.solution[1, i]) < 0.01 && abs(a.solution[4, i] - b.solution[4, i]) < 0.01 &&
    abs(a.masses[1, i] - b.masses[1, i]) < 0.01 && abs(a.omega[i] - b.omega[i]) <= 1e-10 * abs(b.omega[i])

# probes at which the learned-seed, DefaultSeed and MultiSeed roots are not all one root, or a solve failed
function ambiguous(s::Surrogate, ctx, T, mu, xi; p_num, t_num, iterations, ftol, multi_iterations)
    set_judge_mask!(ctx, nothing)
    kw = (xi=xi, p_num=p_num, t_num=t_num, ftol=ftol, ctx=ctx)
    q = solve_batch(T, mu; policy=SEED_SURROGATE, seeds=infer(s, T, mu, xi), iterations=iterations, kw...)
    # MultiSeed probe: Newton only, capped -- it only has to find the competing roots
    m = solve_batch(T, mu; policy=SEED_MULTI, iterations=min(iterations, multi_iterations),
                    trust_region_fallback=false, kw...)
    pr = solve_batch(T, mu; policy=SEED_DEFAULT, iterations=iterations, kw...)
    return [!(q.converged[i] && m.converged[i] && pr.converged[i] && same_root(q, m, i) && same_root(pr, m, i))
            for i in eachindex(T)]
end

bin_of(s::Surrogate, T, mu, xi, nb) = ntuple(d -> clamp(floor(Int, (((T, mu, xi)[d] - s.in_lo[d]) /
                                                                   (s.in_hi[d] - s.in_lo[d])) * nb[d]) + 1,
                                                      1, nb[d]), 3)

# n steps of binary dilation over the 26 neighbours; mask indexed [T, mu, xi]
function dilate(mask::BitArray{3}, n::Integer)
    for _ in 1:n
        g = copy(mask)
        for I in findall(mask), d in CartesianIndices((-1:1, -1:1, -1:1))
            J = I + d
            checkbounds(Bool, g, J) && (g[J] = true)
        end
        mask = g
    end
    return mask
end

"""
    build_judge_mask!(s, ctx, T, mu, xi; bins=(128, 128, 16), refine=8, dilate_bins=1, multi_iterations=12, ...)

The judge's ambiguity mask, learned on probe points: a bin is set where the
learned-seed, plain (DefaultSeed) and MultiSeed (argmin-Omega) roots are not
all one root, or a solve failed; a second pass puts `refine` random probes in
every bin within two bins of a set one; the set bins are dilated over their
26 neighbours.  Stored in `s.judge_mask` and installed in `ctx`.
"""
function build_judge_mask!(s::Surrogate, ctx::Context, T, mu, xi; bins=(128, 128, 16), refine::Integer=8,
                           dilate_bins::Integer=1, multi_iterations::Integer=12, p_num::Integer=64,
                           t_num::Integer=4, iterations::Integer=40, ftol::Real=1e-10, rng=Random.Xoshiro(6))
    kw = (p_num=p_num, t_num=t_num, iterations=iterations, ftol=Float64(ftol), multi_iterations=multi_iterations)
    mask = falses(bins...)
    for i in findall(ambiguous(s, ctx, T, mu, xi; kw...))
        mask[bin_of(s, T[i], mu[i], xi[i], bins)...] = true
    end
    if refine > 0 && any(mask)
        cells = findall(dilate(mask, 2))
        P = Matrix{Float64}(undef, length(cells) * refine, 3)
        k = 0
        for c in cells, _ in 1:refine
            k += 1
            for d in 1:3
                P[k, d] = s.in_lo[d] + ((c[d] - 1) + rand(rng)) / bins[d] * (s.in_hi[d] - s.in_lo[d])
            end
        end
        Tp, mup, xip = P[:, 1], P[:, 2], P[:, 3]
        for i in findall(ambiguous(s, ctx, Tp, mup, xip; kw...))
            mask[bin_of(s, Tp[i], mup[i], xip[i], bins)...] = true
        end
    end
    mask = dilate(mask, dilate_bins)
    s.judge_mask = permutedims(UInt8.(mask), (3, 2, 1))     # the C layout: [T][mu][xi], xi fastest
    set_judge_mask!(ctx, s)
    return s
end

training_points(n, rng) = ([DOMAIN_LO[d] + rand(rng) * (DOMAIN_HI[d] - DOMAIN_LO[d]) for _ in 1:n] for d in 1:3)

# Adam phases (steps, lr) of learned_seed: a step decay, so that the network
# does not end on a late loss spike of a constant rate
const LR_SCHEDULE = ((2000, 3e-3), (2000, 1.5e-3), (2000, 6e-4), (2000, 2e-4))

"""
    learned_seed(ctx; device=0, p_num=64, t_num=4, n_train=200_000, schedule=LR_SCHEDULE, batch=8192,
                 n_probe=1_000_000) -> Surrogate

The learned seed as the scans use it: plain (DefaultSeed) roots at n_train
points of the training box, Adam on them (the (steps, lr) phases of
`schedule`, batch draws seeded 4, 5, ... per phase), and the judge mask from
those points plus n_probe random probes; installed in `ctx` (`device` is the
context's GPU).
"""
function learned_seed(ctx::Context=default_context(); device::Integer=0, p_num::Integer=64, t_num::Integer=4,
                      iterations::Integer=40, ftol::Real=1e-10, n_train::Integer=200_000, schedule=LR_SCHEDULE,
                      batch::Integer=8192, n_probe::Integer=1_000_000, rng=Random.Xoshiro(4))
    saved = launch_cfg(ctx)
    set_launch_cfg!(SETUP_LAUNCH, ctx)
    try
        T, mu, xi = training_points(n_train, rng)
        tr = solve_batch(T, mu; xi=xi, policy=SEED_DEFAULT, p_num=p_num, t_num=t_num, iterations=iterations,
                         ftol=Float64(ftol), ctx=ctx)
        ok = findall(tr.converged)
        s = Surrogate(device)
        fit_output_scaling!(s, tr.solution[:, ok])
        for (i, (steps, lr)) in enumerate(schedule)
            train!(s, T[ok], mu[ok], xi[ok], tr.solution[:, ok]; steps=steps, batch=batch, lr=lr, seed=3 + i)
        end
        Tp, mup, xip = training_points(n_probe, rng)
        build_judge_mask!(s, ctx, vcat(T, Tp), vcat(mu, mup), vcat(xi, xip); p_num=p_num, t_num=t_num,
                          iterations=iterations, ftol=ftol)
        return s
    finally
        set_launch_cfg!(saved, ctx)
    end
end

"""
    SurrogateSeed(model; max_steps=8, judge=:reference)

Seed strategy of the learned seed for `solve` and `solve_many`
(`seed_strategy=SurrogateSeed(model)`): Newton from the model's
seed for at most `max_steps` steps, the DefaultSeed policy where that fails;
inside the model's judge mask `judge=:reference` solves the point with
DefaultSeed unseen, `:omega` keeps the minimum-Omega root among the learned
root and the MultiSeed candidates; `judge=:multiroot3` keeps the minimum-Omega
root of the learned, HADRON and QUARK seeds at every point.
"""
struct SurrogateSeed
    model::Surrogate
    max_steps::Int
    judge::Symbol
end
SurrogateSeed(model::Surrogate; max_steps::Integer=8, judge::Symbol=:reference) =
    SurrogateSeed(model, max_steps, judge)

function solve_learned(mode, T_fm::AbstractVector{Float64}, mu_fm::AbstractVector{Float64}, s::SurrogateSeed;
                       xi::AbstractVector{Float64}=zeros(length(T_fm)), p_num::Integer=64, t_num::Integer=8,
                       iterations::Integer=1000, make_result=SolverResult, ctx::Context=default_context(),
                       kwargs...)
    nameof(typeof(mode)) === :FixedMu || throw(ArgumentError("the HIP backend solves FixedMu() only"))
    s.judge in (:reference, :omega, :multiroot3) || throw(ArgumentError("judge: :reference, :omega or :multiroot3"))
    set_judge_mask!(ctx, s.model)
    seeds = infer(s.model, Vector{Float64}(T_fm), Vector{Float64}(mu_fm), Vector{Float64}(xi))
    r = solve_batch(T_fm, mu_fm; xi=xi, policy=s.judge === :multiroot3 ? SEED_MULTIROOT3 : SEED_SURROGATE,
                    seeds=seeds, p_num=p_num, t_num=t_num, iterations=iterations, surrogate_max_steps=s.max_steps,
                    surrogate_judge=s.judge === :omega ? JUDGE_OMEGA : JUDGE_REFERENCE, ctx=ctx, kwargs...)
    return [make_result(mode, r.converged[i], r.solution[:, i], Tuple(r.solution[:, i]),
                        (mu_fm[i], mu_fm[i], mu_fm[i]), r.omega[i], r.pressure[i], r.rho_norm[i], r.entropy[i],
                        r.energy[i], Tuple(r.masses[:, i]), Int(r.iterations[i]), r.residual_norm[i], xi[i])
            for i in eachindex(T_fm)]
end

end # module PNJLHip
