# Smoke tests of the EqdDevice shim, for a machine with Julia and a gfx950 GPU:
#   julia earthquakediffinitive.jl_amd/julia/runtests.jl
# They mirror checks of the reference's test/fault_response_test.jl and test/bp8_test.jl through
# the shim's own entry points (the C ABI underneath is tested from Python in tests/).
include(joinpath(@__DIR__, "EqdDevice.jl"))
using .EqdDevice
using LinearAlgebra, Test

@testset "EqdDevice" begin
    fe = DeviceFaultElasticity(13, 13, 1.2, 1.2, 2.0, 1.0; l_f=0.6)
    nf = frictional_node_count(fe)
    x2, x3 = fault_grid_axes(fe)
    @test x2[1] ≈ -0.6 && x2[end] ≈ 0.6 && nf == length(x2) * length(x3)
    K = fault_stiffness(fe)
    @test all(<(0), diag(K))
    s2 = [0.01 * sin(3k) for k in 1:nf]
    s3 = [0.02 * cos(2k) for k in 1:nf]
    Δτ2, Δτ3 = shear_traction(fe, s2, s3)
    st = K * vcat(s2, s3)
    @test st[1:nf] ≈ Δτ2 rtol = 1e-8
    @test st[nf+1:2nf] ≈ Δτ3 rtol = 1e-8
    @test solver_report(fe.rs).unconverged == 0
    @test_throws ErrorException DeviceFaultElasticity(13, 13, 1.2, 1.2, 2.0, 1.0; l_f=0.55)
    fj = DeviceFaultElasticity(13, 13, 1.2, 1.2, 2.0, 1.0; l_f=0.6, precond=:jacobi)
    @test fault_stiffness(fj) ≈ K rtol = 1e-8
    minv = jacobi_preconditioner(fj)
    @test length(minv) == fj.Ntot && all(>(0), minv)

    m = build_model(; Δz=100.0, L_fault=800.0, L_normal=800.0)
    u0 = initial_state(m)
    sol = run_bp8(m; tspan=(0.0, 5 * 3600.0), saveat=1800.0)
    @test sol.retcode == :Success && length(sol.t) == 11
    out = saved_outputs(m, sol, [1, m.nf])
    @test size(out.V2) == (2, 11) && all(out.Vmax .> 0)

    e = DeviceEnsemble(m, 4; Q0=fill(m.par.Q0, 4))
    st4 = run_ensemble!(e; tspan=(0.0, 5 * 3600.0), saveat=1800.0)
    @test all(s -> s.retcode == 0, st4)
    s1 = member_solution(e, 1)
    @test s1.t == sol.t
    @test s1.u[end] ≈ sol.u[end] rtol = 1e-6
end
