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
"""
module PNJLHip

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

end # module PNJLHip
