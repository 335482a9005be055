# EqdDevice: the Julia side of libeqd's C ABI (include/eqd.h).
#
# A drop-in for the hot path of EarthquakeDiffinitive.jl: the same function names and
# argument meanings as ElasticitySplitNode (CGSolver, split_node_solve, duplicate,
# merge_stats!, solver_report), FaultResponse (FaultElasticity, shear_traction,
# fault_stiffness, fault_stiffness_toeplitz) and BP8 (build_model, evaluate!, rhs!,
# run_bp8), executed on a gfx950 GPU. Arrays cross the boundary as plain pointers in
# Julia's column-major Float64 layout, so `pointer(K)` is passed directly.
#
# Errors come back as an eqd_status; `check` rethrows them as ErrorException with the
# library's message, as the reference's `error(...)` calls do. An unconverged CG solve
# is counted and warned about once, not thrown (src/ElasticitySplitNode.jl:386-393).
module EqdDevice

export DeviceCGSolver, DeviceFaultElasticity, split_node_solve, duplicate, merge_stats!, solver_report,
       shear_traction, fault_stiffness, fault_stiffness_toeplitz, frictional_node_count, fault_grid_axes,
       DeviceBP8Model, build_model, initial_state, evaluate!, rhs!, run_bp8, state_length,
       effective_stress_report, saved_outputs, DeviceEnsemble, run_ensemble!, member_solution,
       ensemble_saved_vmax, DeviceServer, add_grid!, submit_columns!, submit_forward!, serve_step!,
       completed_requests, request_result, jacobi_preconditioner, set_ks_method!, ks_product

const libeqd = joinpath(@__DIR__, "..", "libeqd.so")

const EQD_OK = Cint(0)
const EQD_PREC = Dict(:fp64 => Cint(0), :mixed => Cint(1))
const EQD_PRECOND = Dict(:none => Cint(0), :jacobi => Cint(1))
const EQD_ENGINE = Dict(:batched => Cint(0), :plain => Cint(1))
const EQD_STIFFNESS = Dict(:exact => Cint(0), :toeplitz => Cint(1))
const EQD_INJECTION = Dict(:gaussian => Cint(0), :peaceman => Cint(1))
const EQD_KS = Dict(:direct => Cint(0), :fft => Cint(1))

struct GridDesc
    n1::Cint
    n23::Cint
    L_normal::Cdouble
    L_fault::Cdouble
    lambda::Cdouble
    mu::Cdouble
    l_f::Cdouble
end

struct OpInfo
    nf::Int64
    n2f::Int64
    n3f::Int64
    ntot::Int64
    nreduced::Int64
    h1::Cdouble
    h2::Cdouble
end

mutable struct CgOpts
    rtol::Cdouble
    atol::Cdouble
    itmax::Int64
    precond::Cint
    precision::Cint
    engine::Cint
    tile_k::Cint
    max_cols_in_flight::Cint
    reliable_interval::Cint
    graph_iters::Cint
    timing::Cint
end
CgOpts() = CgOpts(0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0)

mutable struct Report
    solves::Int64
    iterations::Int64
    max_iterations::Int64
    unconverged::Int64
    mean_iterations::Cdouble
    precond::Cint
    tile_k::Cint
end
Report() = Report(0, 0, 0, 0, 0.0, 0, 0)

function check(rc::Cint)
    rc == EQD_OK && return nothing
    buf = Vector{UInt8}(undef, 1024)
    ccall((:eqd_last_error, libeqd), Csize_t, (Ptr{UInt8}, Csize_t), buf, length(buf))
    error(unsafe_string(pointer(buf)))
end

# ---------------------------------------------------------------------------------
# ElasticitySplitNode / FaultResponse (src/ElasticitySplitNode.jl:365-447,
# src/FaultResponse.jl:96-393)

mutable struct DeviceCGSolver
    handle::Ptr{Cvoid}
    op::Ptr{Cvoid}
end

mutable struct DeviceFaultElasticity
    op::Ptr{Cvoid}
    rs::DeviceCGSolver
    info::OpInfo
    x2f::Vector{Float64}
    x3f::Vector{Float64}
    Ntot::Int
    grid::GridDesc
end

function _solver(op::Ptr{Cvoid}; rtol=1e-10, atol=0.0, itmax=0, precond::Symbol=:none,
                 precision::Symbol=:mixed, engine::Symbol=:batched, tile_k=0)
    haskey(EQD_PRECOND, precond) || error("precond must be :none or :jacobi, got $precond")
    o = CgOpts()
    ccall((:eqd_cg_opts_default, libeqd), Cvoid, (Ref{CgOpts},), o)
    o.rtol, o.atol, o.itmax = rtol, atol, itmax
    o.precond, o.precision, o.engine, o.tile_k = EQD_PRECOND[precond], EQD_PREC[precision], EQD_ENGINE[engine], tile_k
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:eqd_solver_create, libeqd), Cint, (Ptr{Cvoid}, Ref{CgOpts}, Ref{Ptr{Cvoid}}), op, o, h))
    s = DeviceCGSolver(h[], op)
    finalizer(x -> ccall((:eqd_solver_destroy, libeqd), Cint, (Ptr{Cvoid},), x.handle), s)
    return s
end

"""
    DeviceFaultElasticity(n1, n23, L_normal, L_fault, λ, μ; l_f, solver_kwargs...)

FaultElasticity(g₋, g₊, λ, μ, set; l_f, solver_kwargs...) for the two blocks
[-L_normal,0]×[-L_fault,L_fault]² and [0,L_normal]×[...]² with n1×n23×n23 nodes each.
"""
function DeviceFaultElasticity(n1, n23, L_normal, L_fault, λ, μ; l_f, solver_kwargs...)
    g = GridDesc(n1, n23, L_normal, L_fault, λ, μ, l_f)
    op = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:eqd_operator_create, libeqd), Cint, (Ref{GridDesc}, Ref{Ptr{Cvoid}}), g, op))
    info = Ref{OpInfo}()
    check(ccall((:eqd_operator_info, libeqd), Cint, (Ptr{Cvoid}, Ref{OpInfo}), op[], info))
    x2 = zeros(info[].n2f)
    x3 = zeros(info[].n3f)
    check(ccall((:eqd_fault_grid_axes, libeqd), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), op[], x2, x3))
    fe = DeviceFaultElasticity(op[], _solver(op[]; solver_kwargs...), info[], x2, x3, Int(info[].ntot), g)
    finalizer(x -> ccall((:eqd_operator_destroy, libeqd), Cint, (Ptr{Cvoid},), x.op), fe)
    return fe
end

frictional_node_count(fe::DeviceFaultElasticity) = Int(fe.info.nf)

"""
    jacobi_preconditioner(fe) -> Vector{Float64}

The diagonal of jacobi_preconditioner(A)'s M⁻¹ = diag(A)⁻¹ (src/ElasticitySplitNode.jl:343-357)
for fe's split-node A, with A's zero diagonal entries (the far-field rows) floored to 1.
"""
function jacobi_preconditioner(fe::DeviceFaultElasticity)
    minv = zeros(fe.Ntot)
    check(ccall((:eqd_jacobi_preconditioner, libeqd), Cint, (Ref{GridDesc}, Ptr{Float64}, Int64),
                fe.grid, minv, fe.Ntot))
    return minv
end
fault_grid_axes(fe::DeviceFaultElasticity) = (fe.x2f, fe.x3f)

function split_node_solve(s::DeviceCGSolver, rhs::Vector{Float64})
    u = similar(rhs)
    check(ccall((:eqd_split_node_solve, libeqd), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), s.handle, rhs, u))
    return u
end

function duplicate(s::DeviceCGSolver)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:eqd_solver_duplicate, libeqd), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}), s.handle, h))
    d = DeviceCGSolver(h[], s.op)
    finalizer(x -> ccall((:eqd_solver_destroy, libeqd), Cint, (Ptr{Cvoid},), x.handle), d)
    return d
end

function merge_stats!(into::DeviceCGSolver, from::DeviceCGSolver)
    check(ccall((:eqd_solver_merge_stats, libeqd), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), into.handle, from.handle))
    return into
end

function solver_report(s::DeviceCGSolver)
    r = Report()
    check(ccall((:eqd_solver_report, libeqd), Cint, (Ptr{Cvoid}, Ref{Report}), s.handle, r))
    return (; solves=r.solves, iterations=r.iterations, mean_iterations=r.mean_iterations,
              max_iterations=r.max_iterations, unconverged=r.unconverged,
              precond=r.precond == 0 ? :none : :jacobi, tile_k=r.tile_k)
end

function shear_traction(fe::DeviceFaultElasticity, s2, s3)
    nf = frictional_node_count(fe)
    Δτ2, Δτ3 = zeros(nf), zeros(nf)
    check(ccall((:eqd_shear_traction, libeqd), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                fe.rs.handle, collect(Float64, s2), collect(Float64, s3), Δτ2, Δτ3))
    return Δτ2, Δτ3
end

function _stiffness(fe::DeviceFaultElasticity, mode)
    n = 2 * frictional_node_count(fe)
    K = Matrix{Float64}(undef, n, n)
    check(ccall((:eqd_fault_stiffness, libeqd), Cint, (Ptr{Cvoid}, Cint, Ptr{Cint}, Cint, Ptr{Float64}, Int64),
                fe.rs.handle, mode, C_NULL, 0, K, n))
    return K
end

"fault_stiffness(fe; verbose, threaded) (src/FaultResponse.jl:229): all 2nf columns on the device."
fault_stiffness(fe::DeviceFaultElasticity; verbose=false, threaded=false) = _stiffness(fe, EQD_STIFFNESS[:exact])
"fault_stiffness_toeplitz(fe; verbose) (src/FaultResponse.jl:339)."
fault_stiffness_toeplitz(fe::DeviceFaultElasticity; verbose=false) = _stiffness(fe, EQD_STIFFNESS[:toeplitz])

# ---------------------------------------------------------------------------------
# BP8 (src/BP8.jl:184-470)

struct BP8ParamsC
    μ::Cdouble; ν::Cdouble; ρ::Cdouble; c_s::Cdouble
    a::Cdouble; b::Cdouble; D_RS::Cdouble; V_star::Cdouble; f_star::Cdouble
    σ0::Cdouble; τ_init::Cdouble; V_init::Cdouble; V_zero::Cdouble
    Q0::Cdouble; L_gauss::Cdouble; L_fwid::Cdouble; β::Cdouble; φ::Cdouble; k::Cdouble
    viscosity::Cdouble; α::Cdouble; S_well::Cdouble; r_well::Cdouble
    l_f::Cdouble; Δz::Cdouble; t_off::Cdouble; t_f::Cdouble; σ̄_min::Cdouble
end

"Table 1 defaults, with keyword overrides by field name."
function benchmark_parameters(; kw...)
    p = Ref{BP8ParamsC}()
    ccall((:eqd_bp8_default_params, libeqd), Cvoid, (Ref{BP8ParamsC},), p)
    vals = Dict{Symbol,Float64}(f => getfield(p[], f) for f in fieldnames(BP8ParamsC))
    for (k, v) in kw
        haskey(vals, k) || error("unknown BP8 parameter $k")
        vals[k] = v
    end
    return BP8ParamsC((vals[f] for f in fieldnames(BP8ParamsC))...)
end

mutable struct BP8Stats
    naccept::Int64
    nreject::Int64
    nrhs::Int64
    nsaved::Int64
    retcode::Cint
    ms::Cdouble
end
BP8Stats() = BP8Stats(0, 0, 0, 0, 0, 0.0)

# The same record as an isbits value, for arrays passed to the C ABI (a Vector of the mutable
# struct would hold references, not the contiguous records the library writes).
struct BP8StatsC
    naccept::Int64
    nreject::Int64
    nrhs::Int64
    nsaved::Int64
    retcode::Cint
    ms::Cdouble
end

mutable struct DeviceBP8Model
    handle::Ptr{Cvoid}
    par::BP8ParamsC
    nf::Int
    state_len::Int
    fe::Union{Nothing,DeviceFaultElasticity}
end

"""
    build_model(; par, Δz, L_fault, L_normal, injection=:gaussian, stiffness=:toeplitz, solver_kwargs...)

build_model (src/BP8.jl:184-258) with the K build on the device; K stays in device memory.
"""
function build_model(; par::BP8ParamsC=benchmark_parameters(), Δz=par.Δz, L_fault=3par.l_f,
                     L_normal=2par.l_f, injection=:gaussian, order=4, verbose=false, stiffness=:toeplitz,
                     solver_kwargs...)
    injection ∈ (:gaussian, :peaceman) || error("injection must be :gaussian or :peaceman, got $injection")
    isapprox(par.l_f / Δz, round(par.l_f / Δz); atol=1e-9) ||
        error("Δz=$Δz must divide l_f=$(par.l_f) so the grid has nodes on ±l_f")
    L_fault >= par.l_f || error("L_fault=$L_fault must be at least l_f=$(par.l_f)")
    stiffness ∈ (:exact, :toeplitz) || error("stiffness must be :exact or :toeplitz, got $stiffness")
    order == 4 || error("only SBP order 4 is provided")
    n1 = round(Int, L_normal / Δz) + 1
    n23 = round(Int, 2L_fault / Δz) + 1
    n1 >= 9 || error("L_normal/Δz gives only $n1 points across the fault-normal direction; " *
                     "SBP order 4 needs at least 9. Increase L_normal or decrease Δz.")
    n23 >= 9 || error("2*L_fault/Δz gives only $n23 points along the fault; SBP order 4 needs at least 9.")
    λ = 2par.μ * par.ν / (1 - 2par.ν)
    fe = DeviceFaultElasticity(n1, n23, L_normal, L_fault, λ, par.μ; l_f=par.l_f, solver_kwargs...)
    p = BP8ParamsC((f === :Δz ? Float64(Δz) : getfield(par, f) for f in fieldnames(BP8ParamsC))...)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:eqd_bp8_create_from_solver, libeqd), Cint,
                (Ptr{Cvoid}, Cint, Ref{BP8ParamsC}, Cint, Ref{Ptr{Cvoid}}),
                fe.rs.handle, EQD_STIFFNESS[stiffness], p, EQD_INJECTION[injection], h))
    nf, ns = Ref{Cint}(0), Ref{Cint}(0)
    check(ccall((:eqd_bp8_sizes, libeqd), Cint, (Ptr{Cvoid}, Ref{Cint}, Ref{Cint}), h[], nf, ns))
    m = DeviceBP8Model(h[], p, nf[], ns[], fe)
    finalizer(x -> ccall((:eqd_bp8_destroy, libeqd), Cint, (Ptr{Cvoid},), x.handle), m)
    return m
end

state_length(m::DeviceBP8Model) = m.state_len

function initial_state(m::DeviceBP8Model)
    u = zeros(m.state_len)
    check(ccall((:eqd_bp8_initial_state, libeqd), Cint, (Ptr{Cvoid}, Ptr{Float64}), m.handle, u))
    return u
end

"evaluate!(m, u, t) (src/BP8.jl:355-396) -> (; V2, V3, Vmag, τ2, τ3)"
function evaluate!(m::DeviceBP8Model, u::Vector{Float64}, t)
    V2, V3, Vmag, τ2, τ3 = (zeros(m.nf) for _ in 1:5)
    check(ccall((:eqd_bp8_evaluate, libeqd), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Cdouble, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                 Ptr{Float64}), m.handle, u, t, V2, V3, Vmag, τ2, τ3))
    return (; V2, V3, Vmag, τ2, τ3)
end

"rhs!(du, u, m, t) (src/BP8.jl:403-433)"
function rhs!(du::Vector{Float64}, u::Vector{Float64}, m::DeviceBP8Model, t)
    check(ccall((:eqd_bp8_rhs, libeqd), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cdouble, Ptr{Float64}),
                m.handle, u, t, du))
    return du
end

"""
    set_ks_method!(m, :fft | :direct)

How a `stiffness=:toeplitz` model forms K s in rhs! (src/FaultResponse.jl:378-390 makes K a
function of node separation): the 2D FFT convolution (default beyond 48 x 48 fault nodes) or the
direct correlation. Errors for a dense-K model.
"""
function set_ks_method!(m::DeviceBP8Model, method::Symbol)
    haskey(EQD_KS, method) || error("K s method must be :fft or :direct, got $method")
    check(ccall((:eqd_bp8_set_ks_method, libeqd), Cint, (Ptr{Cvoid}, Cint), m.handle, EQD_KS[method]))
    return m
end

"K s = [τ2; τ3] for a slip vector [s2; s3] with the model's K s method."
function ks_product(m::DeviceBP8Model, s::Vector{Float64})
    length(s) == 2m.nf || error("slip must have $(2m.nf) entries")
    out = zeros(2m.nf)
    check(ccall((:eqd_bp8_ks, libeqd), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), m.handle, s, out))
    return out
end

"""
    run_bp8(m; tspan=(0.0, m.par.t_f), reltol=1e-8, saveat=3600.0, maxiters=100000, alg=:tsit5)
        -> (; t, u, stats)

run_bp8 (src/BP8.jl:442-470): Tsit5 with OrdinaryDiffEq's defaults, stepped on the device;
alg=:imex runs the implicit-local integrator (L-stable SDIRK4 stages solved by per-node 3 x 3
Newton blocks) for the stiff Peaceman-well variant.
"""
function run_bp8(m::DeviceBP8Model; tspan=(0.0, m.par.t_f), reltol=1e-8, saveat=3600.0, verbose=false,
                 maxiters=100000, alg::Symbol=:tsit5)
    alg in (:tsit5, :imex) || error("alg must be :tsit5 or :imex, got $alg")
    t0, tf = Float64.(tspan)
    nmax = floor(Int, (tf - t0) / saveat + 1e-9) + 2
    ts = zeros(nmax)
    us = zeros(m.state_len, nmax)     # column per save == the C ABI's row per save
    st = BP8Stats()
    if alg === :tsit5
        check(ccall((:eqd_bp8_integrate, libeqd), Cint,
                    (Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cdouble, Int64, Int64, Ptr{Float64}, Ptr{Float64},
                     Ref{BP8Stats}), m.handle, t0, tf, saveat, reltol, maxiters, nmax, ts, us, st))
    else
        check(ccall((:eqd_bp8_integrate_imex, libeqd), Cint,
                    (Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cdouble, Int64, Int64, Ptr{Float64}, Ptr{Float64},
                     Ref{BP8Stats}), m.handle, t0, tf, saveat, reltol, maxiters, nmax, ts, us, st))
    end
    n = st.nsaved
    return (; t=ts[1:n], u=[us[:, j] for j in 1:n], stats=(; naccept=st.naccept, nreject=st.nreject, nf=st.nrhs),
              retcode=st.retcode == 0 ? :Success : :MaxIters)
end

function effective_stress_report(m::DeviceBP8Model)
    lo, hits = Ref{Cdouble}(0), Ref{Int64}(0)
    mask = zeros(UInt8, m.nf)
    check(ccall((:eqd_bp8_stress_report, libeqd), Cint, (Ptr{Cvoid}, Ref{Cdouble}, Ref{Int64}, Ptr{UInt8}),
                m.handle, lo, hits, mask))
    nodes = count(!iszero, mask)
    return (; σ̄_lowest=lo[], floor_hits=hits[], floor=m.par.σ̄_min, bound=hits[] > 0, nodes,
              fraction=nodes / m.nf)
end

"""
    saved_outputs(m, sol, idx) -> (; V2, V3, τ2, τ3, q2, q3, Vmax, moment_rate)

The fields write_outputs (src/BP8.jl:605-642) recomputes with evaluate! at every saved time,
evaluated on the device in one pass over the saves of the last run_bp8: slip rates, shear
stresses and Darcy velocities at the Ω_f nodes `idx` (1-based linear, x2 fastest) as
`length(idx) × nsaves` matrices, and max |V| and the moment rate per save. The reference's
writers (write_time_series, write_global, write_profiles) take these in place of their
evaluate! loop.
"""
function saved_outputs(m::DeviceBP8Model, sol, idx::AbstractVector{<:Integer})
    ns = length(sol.t)
    ci = Cint.(idx .- 1)
    f = zeros(6, length(ci), ns)
    g = zeros(2, ns)
    check(ccall((:eqd_bp8_saved_outputs, libeqd), Cint,
                (Ptr{Cvoid}, Int64, Ptr{Cint}, Cint, Ptr{Float64}, Ptr{Float64}),
                m.handle, ns, ci, length(ci), f, g))
    return (; V2=f[1, :, :], V3=f[2, :, :], τ2=f[3, :, :], τ3=f[4, :, :], q2=f[5, :, :], q3=f[6, :, :],
              Vmax=g[1, :], moment_rate=g[2, :])
end

# ---------------------------------------------------------------------------------------
# Trajectory ensembles (no counterpart in the reference, which runs one trajectory per
# call): M members of run_bp8 on one model's K, each with its own Q0, a, σ0 (and reltol);
# K S over all members is one FFT convolution per stage for a :toeplitz model, else one FP64
# MFMA GEMM.

mutable struct DeviceEnsemble
    handle::Ptr{Cvoid}
    model::DeviceBP8Model
    members::Int
    stats::Vector{BP8StatsC}
end

function DeviceEnsemble(m::DeviceBP8Model, members::Integer; Q0=nothing, a=nothing, σ0=nothing)
    arr(v) = v === nothing ? C_NULL : (length(v) == members ? Float64.(collect(v)) :
                                        error("member parameter arrays need one value per member"))
    q, aa, sg = arr(Q0), arr(a), arr(σ0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve q aa sg begin
        check(ccall((:eqd_bp8_ensemble_create, libeqd), Cint,
                    (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ref{Ptr{Cvoid}}),
                    m.handle, members, q, aa, sg, h))
    end
    e = DeviceEnsemble(h[], m, members, BP8StatsC[])
    finalizer(x -> ccall((:eqd_bp8_ensemble_destroy, libeqd), Cint, (Ptr{Cvoid},), x.handle), e)
    return e
end

"""Integrates every member (run_bp8's controller per member); returns the members' stats.
alg=:imex runs the implicit-local integrator over the members (eqd_bp8_ensemble_integrate_imex)."""
function run_ensemble!(e::DeviceEnsemble; tspan=(0.0, e.model.par.t_f), reltol=1e-8, member_reltol=nothing,
                       saveat=3600.0, maxiters=100000, alg::Symbol=:tsit5)
    alg in (:tsit5, :imex) || error("alg must be :tsit5 or :imex, got $alg")
    t0, tf = Float64.(tspan)
    nmax = floor(Int, (tf - t0) / saveat + 1e-9) + 2
    st = Vector{BP8StatsC}(undef, e.members)
    rt = member_reltol === nothing ? C_NULL : Float64.(collect(member_reltol))
    ms = Ref{Cdouble}(0)
    GC.@preserve rt begin
        if alg === :tsit5
            check(ccall((:eqd_bp8_ensemble_integrate, libeqd), Cint,
                        (Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cdouble, Ptr{Cdouble}, Int64, Int64, Ptr{BP8StatsC},
                         Ref{Cdouble}), e.handle, t0, tf, saveat, reltol, rt, maxiters, nmax, st, ms))
        else
            check(ccall((:eqd_bp8_ensemble_integrate_imex, libeqd), Cint,
                        (Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cdouble, Ptr{Cdouble}, Int64, Int64, Ptr{BP8StatsC},
                         Ref{Cdouble}), e.handle, t0, tf, saveat, reltol, rt, maxiters, nmax, st, ms))
        end
    end
    e.stats = st
    return st
end

"Member j's (1-based) saves as (; t, u) like run_bp8 returns."
function member_solution(e::DeviceEnsemble, j::Integer)
    n = Ref{Int64}(0)
    check(ccall((:eqd_bp8_ensemble_saved, libeqd), Cint, (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Ref{Int64}),
                e.handle, j - 1, C_NULL, C_NULL, n))
    ts = zeros(n[])
    us = zeros(e.model.state_len, n[])
    check(ccall((:eqd_bp8_ensemble_saved, libeqd), Cint, (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Ref{Int64}),
                e.handle, j - 1, ts, us, n))
    return (; t=ts, u=[us[:, k] for k in 1:n[]])
end

"max |V| at the first n saves of every member (n × members), one GEMM per save."
function ensemble_saved_vmax(e::DeviceEnsemble, n::Integer)
    v = zeros(n, e.members)
    check(ccall((:eqd_bp8_ensemble_saved_vmax, libeqd), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Int64, Ptr{Cdouble}),
                e.handle, v, n, C_NULL))
    return v
end

# ---------------------------------------------------------------------------------------
# Serving loop (no counterpart in the reference, whose K build is a batch job): stiffness
# columns and forward solves on resident grids, continuously batched.

mutable struct DeviceServer
    handle::Ptr{Cvoid}
end

function DeviceServer(; rtol=1e-10, precision::Symbol=:mixed, hbm_budget_bytes::Integer=0)
    o = CgOpts()
    ccall((:eqd_cg_opts_default, libeqd), Cvoid, (Ref{CgOpts},), o)
    o.rtol = rtol
    o.precision = EQD_PREC[precision]
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:eqd_server_create, libeqd), Cint, (Ref{CgOpts}, Csize_t, Ref{Ptr{Cvoid}}),
                o, hbm_budget_bytes, h))
    sv = DeviceServer(h[])
    finalizer(x -> ccall((:eqd_server_destroy, libeqd), Cint, (Ptr{Cvoid},), x.handle), sv)
    return sv
end

"Makes a grid resident; returns (grid id, lanes)."
function add_grid!(sv::DeviceServer, n1, n23, L_normal, L_fault, λ, μ; l_f, max_columns=0)
    g = GridDesc(n1, n23, L_normal, L_fault, λ, μ, l_f)
    id, slots = Ref{Cint}(0), Ref{Cint}(0)
    check(ccall((:eqd_server_add_grid, libeqd), Cint, (Ptr{Cvoid}, Ref{GridDesc}, Cint, Ref{Cint}, Ref{Cint}),
                sv.handle, g, max_columns, id, slots))
    return id[], slots[]
end

"Queues the unit-slip columns `cols` (1-based) of K on a grid; returns the request id."
function submit_columns!(sv::DeviceServer, grid::Integer, cols::AbstractVector{<:Integer})
    c = Cint.(cols .- 1)
    rid = Ref{Int64}(0)
    check(ccall((:eqd_server_submit_columns, libeqd), Cint, (Ptr{Cvoid}, Cint, Ptr{Cint}, Cint, Ref{Int64}),
                sv.handle, grid, c, length(c), rid))
    return rid[]
end

"Queues shear_traction(fe, s2, s3) on a grid; returns the request id."
function submit_forward!(sv::DeviceServer, grid::Integer, s2::Vector{Float64}, s3::Vector{Float64})
    rid = Ref{Int64}(0)
    check(ccall((:eqd_server_submit_forward, libeqd), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Ref{Int64}), sv.handle, grid, s2, s3, rid))
    return rid[]
end

"One step of every grid; returns (lanes busy, requests queued)."
function serve_step!(sv::DeviceServer)
    busy, queued = Ref{Int64}(0), Ref{Int64}(0)
    check(ccall((:eqd_server_step, libeqd), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}), sv.handle, busy, queued))
    return busy[], queued[]
end

function completed_requests(sv::DeviceServer; max=1024)
    ids = zeros(Int64, max)
    n = Ref{Cint}(0)
    check(ccall((:eqd_server_completed, libeqd), Cint, (Ptr{Cvoid}, Ptr{Int64}, Cint, Ref{Cint}),
                sv.handle, ids, max, n))
    return ids[1:n[]]
end

"Result of a completed request: [Δτ2; Δτ3] per column (2nf × ncols), iterations, unconverged."
function request_result(sv::DeviceServer, id::Integer, nf::Integer, ncols::Integer)
    out = zeros(2nf, ncols)
    iters = zeros(Cint, ncols)
    unc = Ref{Cint}(0)
    check(ccall((:eqd_server_result, libeqd), Cint, (Ptr{Cvoid}, Int64, Ptr{Cdouble}, Int64, Ptr{Cint}, Ref{Cint}),
                sv.handle, id, out, 2nf, iters, unc))
    return out, iters, unc[]
end

end # module
