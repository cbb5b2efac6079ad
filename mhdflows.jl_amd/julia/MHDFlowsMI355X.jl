"""
    MHDFlowsMI355X

Julia binding a maintainer adds to MHDFlows.jl so that `Problem(MI355X(); ...)` runs the
3D HD/MHD right-hand side and RK4/LSRK54 stepper on an AMD Instinct MI355X through the C ABI of
`libmhdflows_amd.so` (`include/mhdflows_amd.h`).  Source only: Julia is not installed where this
library was built and tested; the same ABI is exercised from Python (`mhdflows_amd`).

Method map (reference → this file):
  Problem(dev; ...)                       src/pgen.jl:64-127        → Problem(::MI355X; ...)
  SetUpProblemIC!(prob; ux..bz)           src/utils/IC.jl:41-109    → SetUpProblemIC!(::MI355XProblem; ...)
  stepforward!(sol, clock, ts, ...)       src/timestepper/timestepper.jl:4-6 → stepforward!(::MI355XProblem)
  TimeIntegrator!(prob, t₀, N₀; ...)      src/integrator.jl:31-156  → TimeIntegrator!(::MI355XProblem, ...)
  MHDcalcN!(N, sol, t, ...)               src/pgen.jl:153-162       → MHDcalcN!(N::DevPtr, sol::DevPtr, t, ..., grid::MI355XGrid)
  MHDcalcN_advection!(N, sol, t, ...)     src/Solver/MHDSolver.jl:330-351
  UᵢUpdate!(N, ...; direction)            src/Solver/MHDSolver.jl:27-103
  BᵢUpdate!(N, ...; direction)            src/Solver/MHDSolver.jl:106-177
  getCFL!(prob, t_diff; Coef)             src/integrator.jl:158-198
  ProbDiagnostic(prob)                    src/utils/UserInterface.jl:65-86
"""
module MHDFlowsMI355X

using FourierFlows
import MHDFlows

export MI355X

const libmhd = joinpath(@__DIR__, "..", "mhdflows_amd", "libmhdflows_amd.so")

"The MI355X device: state lives in HBM owned by the HIP library."
struct MI355X <: FourierFlows.Device end

# mirror of `mhd_config` (field order and types must match include/mhdflows_amd.h)
struct MhdConfig
    nx::Int32; ny::Int32; nz::Int32
    Lx::Float64; Ly::Float64; Lz::Float64
    nu::Float64; eta::Float64
    n_nu::Int32; n_eta::Int32
    precision::Int32; physics::Int32; stepper::Int32; device::Int32
    stream::Ptr{Cvoid}
    hbm_budget::UInt64
    dt::Float64
end

struct MhdDiag
    E_kin::Float64; E_mag::Float64; H_kin::Float64; H_mag::Float64; H_cross::Float64
    max_div_b::Float64; max_div_u::Float64; has_nan::Int32; pad::Int32
end

function check(rc::Integer)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:mhd_last_error, libmhd), Cstring, ()))
    error(msg)                                # the reference raises error(...) as well
end

"Grid handle passed where the reference passes `grid` (carries the library handle)."
struct MI355XGrid{T}
    h::Ptr{Cvoid}
    nx::Int; ny::Int; nz::Int; nkr::Int
    Lx::T; Ly::T; Lz::T
end

mutable struct MI355XProblem{T}
    h::Ptr{Cvoid}
    cfg::MhdConfig
    grid::MI355XGrid{T}
    nf::Int
    usr_func::Vector{Function}
end

function MHDFlows.Problem(::MI355X; nx=64, ny=nx, nz=nx, Lx=2π, Ly=Lx, Lz=Lx, ν=0.0, nν=0, η=0.0, nη=0,
                          B_field=false, stepper="RK4", T=Float32, dt=0.0, aliased_fraction=1/3,
                          calcF=MHDFlows.nothingfunction, device=0, hbm_budget=0, kwargs...)
    # aliased_fraction and nη are accepted and ignored, as in src/pgen.jl:91,107 and :114-116
    cfg = MhdConfig(nx, ny, nz, Lx, Ly, Lz, ν, B_field ? η : 0.0, nν, 0,
                    T == Float32 ? 0 : 1, B_field ? 1 : 0, stepper == "RK4" ? 0 : 1, device,
                    C_NULL, hbm_budget, dt)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:mhd_create, libmhd), Int32, (Ref{MhdConfig}, Ref{Ptr{Cvoid}}), cfg, h))
    grid = MI355XGrid{T}(h[], nx, ny, nz, div(nx, 2) + 1, T(Lx), T(Ly), T(Lz))
    prob = MI355XProblem{T}(h[], cfg, grid, B_field ? 6 : 3, Function[])
    finalizer(p -> ccall((:mhd_destroy, libmhd), Int32, (Ptr{Cvoid},), p.h), prob)
    return prob
end

function MHDFlows.SetUpProblemIC!(prob::MI355XProblem{T}; ux=[], uy=[], uz=[], bx=[], by=[], bz=[],
                                  kwargs...) where T
    fields = prob.nf == 6 ? (ux, uy, uz, bx, by, bz) : (ux, uy, uz)
    arrays = [f == [] ? nothing : convert(Array{T,3}, f) for f in fields]
    ptrs = Ptr{Cvoid}[a === nothing ? C_NULL : pointer(a) for a in arrays]
    GC.@preserve arrays check(ccall((:mhd_set_ic_real, libmhd), Int32, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}),
                                    prob.h, ptrs))
    return nothing
end

"Copy `prob.sol` (Complex{T}[nkr, ny, nz, nf], the reference layout) to the host."
function solution(prob::MI355XProblem{T}) where T
    g = prob.grid
    out = Array{Complex{T},4}(undef, g.nkr, g.ny, g.nz, prob.nf)
    check(ccall((:mhd_get_sol, libmhd), Int32, (Ptr{Cvoid}, Ptr{Cvoid}), prob.h, out))
    return out
end

MHDFlows.stepforward!(prob::MI355XProblem) =
    check(ccall((:mhd_step, libmhd), Int32, (Ptr{Cvoid}, Int32), prob.h, 1))

function MHDFlows.TimeIntegrator!(prob::MI355XProblem, t₀::Number, N₀::Int; usr_dt=0.0, CFL_Coef=0.25,
                                  CFL_function=MHDFlows.nothingfunction, save=false, kwargs...)
    if save
        error("Save Function Turned ON but save_loc/filename/dump_dt is not declared!\n")
    end
    if CFL_function != MHDFlows.nothingfunction && usr_dt > 0.0
        error("User define both CFL_function and usr_dt")
    end
    steps = Ref{Int64}(0)
    check(ccall((:mhd_integrate, libmhd), Int32,
                (Ptr{Cvoid}, Float64, Int64, Float64, Float64, Ref{Int64}),
                prob.h, Float64(t₀), N₀, Float64(usr_dt), Float64(CFL_Coef), steps))
    return nothing
end

function MHDFlows.getCFL!(prob::MI355XProblem, t_diff; Coef=0.3)
    dt = Ref{Float64}(0.0)
    check(ccall((:mhd_cfl_dt, libmhd), Int32, (Ptr{Cvoid}, Float64, Ref{Float64}), prob.h, Coef, dt))
    return dt[]
end

# Dealiased steps (default on): every step of a call but the last updates kept modes only;
# the state after the call is the same as with every step computed in full.
set_step_pruning!(prob::MI355XProblem, enable::Bool) =
    check(ccall((:mhd_set_step_pruning, libmhd), Int32, (Ptr{Cvoid}, Int32), prob.h, Int32(enable)))

# Operator-level methods: N and sol are device arrays (ROCArray) in the reference layout;
# only their device pointers cross the boundary.
devptr(x) = Ptr{Cvoid}(UInt(pointer(x)))

MHDFlows.MHDcalcN!(N, sol, t, clock, vars, params, grid::MI355XGrid) =
    check(ccall((:mhd_calcN, libmhd), Int32, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Float64),
                grid.h, devptr(N), devptr(sol), Float64(t)))

MHDFlows.MHDSolver.MHDcalcN_advection!(N, sol, t, clock, vars, params, grid::MI355XGrid) =
    check(ccall((:mhd_advection, libmhd), Int32, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Float64),
                grid.h, devptr(N), devptr(sol), Float64(t)))

const DIRS = Dict("x" => 0, "y" => 1, "z" => 2)

function MHDFlows.MHDSolver.UᵢUpdate!(N, sol, t, clock, vars, params, grid::MI355XGrid; direction="x")
    haskey(DIRS, direction) || error("Warning : Unknown direction is declerad")
    check(ccall((:mhd_u_update, libmhd), Int32, (Ptr{Cvoid}, Ptr{Cvoid}, Int32),
                grid.h, devptr(N), DIRS[direction]))
end

function MHDFlows.MHDSolver.BᵢUpdate!(N, sol, t, clock, vars, params, grid::MI355XGrid; direction="x")
    haskey(DIRS, direction) || error("Warning : Unknown direction is declerad")
    check(ccall((:mhd_b_update, libmhd), Int32, (Ptr{Cvoid}, Ptr{Cvoid}, Int32),
                grid.h, devptr(N), DIRS[direction]))
end

function MHDFlows.ProbDiagnostic(prob::MI355XProblem)
    d = Ref{MhdDiag}()
    check(ccall((:mhd_diagnostics, libmhd), Int32, (Ptr{Cvoid}, Ref{MhdDiag}), prob.h, d))
    isnan(d[].E_kin) && error("detected NaN! Quit the simulation right now.")
    KE = round(d[].E_kin, sigdigits=3)
    return prob.nf == 6 ? (KE, round(d[].E_mag, sigdigits=3)) : KE
end

end # module
