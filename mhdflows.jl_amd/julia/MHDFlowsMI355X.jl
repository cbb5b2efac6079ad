"""
    MHDFlowsMI355X

Julia binding a maintainer adds to MHDFlows.jl so that `Problem(MI355X(); ...)` runs the
3D HD/MHD right-hand side and RK4/LSRK54 stepper on an AMD Instinct MI355X through the C ABI of
`libmhdflows_amd.so` (`include/mhdflows_amd.h`).  Source only: Julia is not installed where this
library was built and tested; the same ABI is exercised from Python (`mhdflows_amd`).

Method map (reference → this file):
  Problem(dev; ...)                       src/pgen.jl:64-127        → Problem(::MI355X; ...)
  prob.sol / clock / vars / grid / params / flag / usr_func  (FourierFlows.Problem fields and
                                          MHDFlows' additions, src/pgen.jl:117-127) → properties
  SetUpProblemIC!(prob; ux..bz)           src/utils/IC.jl:41-109    → SetUpProblemIC!(::MI355XProblem; ...)
  stepforward!(sol, clock, ts, ...)       src/timestepper/timestepper.jl:4-6 → stepforward!(::MI355XProblem)
  TimeIntegrator!(prob, t₀, N₀; ...)      src/integrator.jl:31-156  → TimeIntegrator!(::MI355XProblem, ...)
  getCFL!(prob, t_diff; Coef)             src/integrator.jl:158-198
  savefile / Restart!                     src/integrator.jl:208-288
  readMHDFlows(FileName)                  src/utils/IC.jl:245-257
  MHDcalcN!(N, sol, t, ...)               src/pgen.jl:153-162       → MHDcalcN!(N, sol, ..., grid::MI355XGrid)
  MHDcalcN_advection!(N, sol, t, ...)     src/Solver/MHDSolver.jl:330-351
  UᵢUpdate!(N, ...; direction)            src/Solver/MHDSolver.jl:27-103
  BᵢUpdate!(N, ...; direction)            src/Solver/MHDSolver.jl:106-177
  ProbDiagnostic(prob)                    src/utils/UserInterface.jl:65-86
  SetUpN97!, SetUpND!, SetUpFk            src/pgen/TaylorGreenDynamo.jl:18-33, NegativeDamping.jl:16-22,
                                          A99ForceDriving.jl:93-127 (the forcing runs inside the device step)
  spectralline, ScaleDecomposition        src/utils/MHDAnalysis.jl:17-82,236-255 (of prob.vars, on the device)
  SF₂1D / SFr(...; "Global")              src/utils/TurbStatTool.jl:88-160 → SFr(prob, mode)
  DivVCorrection!, DivBCorrection!        src/Solver/VPSolver.jl:60-139
  VP_method, params.χ/U₀/B₀               src/Solver/VPSolver.jl:21-56 → Problem(...; VP_method), SetUpVP!
  Curl, Div, ∂i, LaplaceSolver, VectorPotential  src/utils/VectorCalculus.jl, MHDAnalysis.jl:129-180
                                          → the same names with the problem first (its grid, on the device)
"""
module MHDFlowsMI355X

using FourierFlows
import MHDFlows

export MI355X, set_step_pruning!, autotune!, setsol!, SetUpN97!, SetUpND!, SetUpFk, spectra, SFr, SetUpVP!

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
    dx::T; dy::T; dz::T
    x::Vector{T}; y::Vector{T}; z::Vector{T}   # ThreeDGrid points, origin -L/2
end

Base.eltype(::MI355XGrid{T}) where T = T

"`prob.clock`: t, dt and step read from and written to the device clock."
struct MI355XClock
    h::Ptr{Cvoid}
end

function Base.getproperty(c::MI355XClock, s::Symbol)
    s === :h && return getfield(c, :h)
    t = Ref{Float64}(0); dt = Ref{Float64}(0); st = Ref{Int64}(0)
    check(ccall((:mhd_get_clock, libmhd), Int32, (Ptr{Cvoid}, Ref{Float64}, Ref{Float64}, Ref{Int64}),
                getfield(c, :h), t, dt, st))
    s === :t && return t[]
    s === :dt && return dt[]
    s === :step && return Int(st[])
    error("clock has no field $s")
end

function Base.setproperty!(c::MI355XClock, s::Symbol, v)
    h = getfield(c, :h)
    t = Ref{Float64}(0); dt = Ref{Float64}(0); st = Ref{Int64}(0)
    check(ccall((:mhd_get_clock, libmhd), Int32, (Ptr{Cvoid}, Ref{Float64}, Ref{Float64}, Ref{Int64}), h, t, dt, st))
    s === :t ? (t[] = v) : s === :dt ? (dt[] = v) : s === :step ? (st[] = v) : error("clock has no field $s")
    check(ccall((:mhd_set_clock, libmhd), Int32, (Ptr{Cvoid}, Float64, Float64, Int64), h, t[], dt[], st[]))
    return v
end

const FIELDS = (:ux, :uy, :uz, :bx, :by, :bz)

"`prob.vars`: the real fields, copied to host `Array{T,3}` (nx, ny, nz) on access."
struct MI355XVars{T}
    h::Ptr{Cvoid}
    nx::Int; ny::Int; nz::Int; nf::Int
end

function Base.getproperty(v::MI355XVars{T}, s::Symbol) where T
    s in fieldnames(MI355XVars) && return getfield(v, s)
    f = findfirst(==(s), FIELDS)
    (f === nothing || f > getfield(v, :nf)) && error("vars has no field $s")
    out = Array{T,3}(undef, getfield(v, :nx), getfield(v, :ny), getfield(v, :nz))
    check(ccall((:mhd_get_vars, libmhd), Int32, (Ptr{Cvoid}, Int32, Ptr{Cvoid}), getfield(v, :h), f - 1, out))
    return out
end

"`prob.params`: the reference's MHDParams fields that host code reads (src/Solver/MHDParams.jl)."
struct MI355XParams
    ν::Float64; η::Float64; nν::Int; nη::Int
    ux_ind::Int; uy_ind::Int; uz_ind::Int; bx_ind::Int; by_ind::Int; bz_ind::Int
    calcF!::Function
end

"`prob.flag`: b (MHD), e (EMHD), vp, c (compressible), s (shear) — b and vp can be set here."
struct MI355XFlag
    b::Bool; e::Bool; vp::Bool; c::Bool; s::Bool
end

mutable struct MI355XProblem{T}
    h::Ptr{Cvoid}
    cfg::MhdConfig
    grid::MI355XGrid{T}
    clock::MI355XClock
    vars::MI355XVars{T}
    params::MI355XParams
    flag::MI355XFlag
    nf::Int
    usr_func::Vector{Function}
end

# `prob.sol` is the spectral state Complex{T}[nkr, ny, nz, nf] copied to the host; `setsol!`
# writes one back
function Base.getproperty(p::MI355XProblem, s::Symbol)
    s === :sol && return solution(p)
    return getfield(p, s)
end

# the device-step forcing modules a calcF may name (their set-up functions below upload the data)
N97ForceDriving!(args...) = error("N97ForceDriving! runs inside the device step; call SetUpN97! after Problem")
NDForceDriving!(args...) = error("NDForceDriving! runs inside the device step; call SetUpND! after Problem")
A99ForceDriving!(args...) = error("A99ForceDriving! runs inside the device step; call SetUpFk after Problem")
const DEVICE_FORCING = (N97ForceDriving!, NDForceDriving!, A99ForceDriving!)

function MHDFlows.Problem(::MI355X; nx=64, ny=nx, nz=nx, Lx=2π, Ly=Lx, Lz=Lx, ν=0.0, nν=0, η=0.0, nη=0,
                          B_field=false, stepper="RK4", T=Float32, dt=0.0, aliased_fraction=1/3,
                          calcF=MHDFlows.nothingfunction, usr_func=[], VP_method=false, device=0, hbm_budget=0,
                          kwargs...)
    # aliased_fraction and nη are accepted and ignored, as in src/pgen.jl:91,107 and :114-116
    if calcF !== MHDFlows.nothingfunction && !(calcF in DEVICE_FORCING)
        # addforcing! calls calcF! inside every right-hand side (src/pgen.jl:231-234): a Julia
        # function cannot run there, so it is refused instead of silently dropped
        error("a user calcF! cannot run inside the device step; use N97ForceDriving!, " *
              "NDForceDriving! or A99ForceDriving! with their set-up function")
    end
    cfg = MhdConfig(nx, ny, nz, Lx, Ly, Lz, ν, B_field ? η : 0.0, nν, 0,
                    T == Float32 ? 0 : 1, B_field ? 1 : 0, stepper == "RK4" ? 0 : 1, device,
                    C_NULL, hbm_budget, dt)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:mhd_create, libmhd), Int32, (Ref{MhdConfig}, Ref{Ptr{Cvoid}}), cfg, h))
    dx, dy, dz = T(Lx / nx), T(Ly / ny), T(Lz / nz)
    pts(L, d, n) = collect(range(T(-L / 2), step=d, length=n))
    grid = MI355XGrid{T}(h[], nx, ny, nz, div(nx, 2) + 1, T(Lx), T(Ly), T(Lz), dx, dy, dz,
                         pts(Lx, dx, nx), pts(Ly, dy, ny), pts(Lz, dz, nz))
    nf = B_field ? 6 : 3
    params = MI355XParams(ν, η, nν, nη, 1, 2, 3, 4, 5, 6, calcF)
    flag = MI355XFlag(B_field, false, VP_method, false, false)
    # MHDParams_VP / HDParams_VP: chi, U0 (and B0) start as zero fields (datastructure.jl:79-95)
    VP_method && check(ccall((:mhd_set_vp, libmhd), Int32, (Ptr{Cvoid}, Int32, Ptr{Cvoid}, Ptr{Ptr{Cvoid}},
                                                          Ptr{Ptr{Cvoid}}), h[], 1, C_NULL, C_NULL, C_NULL))
    prob = MI355XProblem{T}(h[], cfg, grid, MI355XClock(h[]), MI355XVars{T}(h[], nx, ny, nz, nf),
                            params, flag, nf, Function[f for f in usr_func])
    finalizer(p -> ccall((:mhd_destroy, libmhd), Int32, (Ptr{Cvoid},), getfield(p, :h)), prob)
    return prob
end

function MHDFlows.SetUpProblemIC!(prob::MI355XProblem{T}; ux=[], uy=[], uz=[], bx=[], by=[], bz=[],
                                  U₀x=[], U₀y=[], U₀z=[], B₀x=[], B₀y=[], B₀z=[], kwargs...) where T
    if !all(isempty, (U₀x, U₀y, U₀z, B₀x, B₀y, B₀z))          # IC.jl:93-105
        prob.flag.vp || error("U₀/B₀ fields need a VP_method problem")
        SetUpVP!(prob; U₀x=U₀x, U₀y=U₀y, U₀z=U₀z, B₀x=B₀x, B₀y=B₀y, B₀z=B₀z)
    end
    fields = prob.nf == 6 ? (ux, uy, uz, bx, by, bz) : (ux, uy, uz)
    arrays = [f == [] ? nothing : convert(Array{T,3}, f) for f in fields]
    ptrs = Ptr{Cvoid}[a === nothing ? C_NULL : pointer(a) for a in arrays]
    GC.@preserve arrays check(ccall((:mhd_set_ic_real, libmhd), Int32, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}),
                                    prob.h, ptrs))
    return nothing
end

"""Volume-penalisation fields of a VP_method problem: `χ` (what the reference's user copies into
`prob.params.χ`, e.g. `Cylindrical_Mask_Function(prob.grid)`), `U₀x..z`, `B₀x..z`; a field left
empty keeps its value.  Every right-hand side then adds VP_UᵢUpdate!/VP_BᵢUpdate!."""
function SetUpVP!(prob::MI355XProblem{T}; χ=[], U₀x=[], U₀y=[], U₀z=[], B₀x=[], B₀y=[], B₀z=[]) where T
    conv(a) = isempty(a) ? nothing : convert(Array{T,3}, a)
    arrays = [conv(a) for a in (χ, U₀x, U₀y, U₀z, B₀x, B₀y, B₀z)]
    p(a) = a === nothing ? C_NULL : pointer(a)
    U = Ptr{Cvoid}[p(a) for a in arrays[2:4]]
    B = Ptr{Cvoid}[p(a) for a in arrays[5:7]]
    GC.@preserve arrays check(ccall((:mhd_set_vp, libmhd), Int32,
                                    (Ptr{Cvoid}, Int32, Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}),
                                    prob.h, 1, p(arrays[1]), U, B))
end

"Copy `prob.sol` (Complex{T}[nkr, ny, nz, nf], the reference layout) to the host."
function solution(prob::MI355XProblem{T}) where T
    g = getfield(prob, :grid)
    out = Array{Complex{T},4}(undef, g.nkr, g.ny, g.nz, getfield(prob, :nf))
    check(ccall((:mhd_get_sol, libmhd), Int32, (Ptr{Cvoid}, Ptr{Cvoid}), getfield(prob, :h), out))
    return out
end

"Write a host spectral state (the layout of `prob.sol`) to the device; vars follow."
function setsol!(prob::MI355XProblem{T}, sol::AbstractArray) where T
    A = convert(Array{Complex{T},4}, sol)
    GC.@preserve A check(ccall((:mhd_set_sol, libmhd), Int32, (Ptr{Cvoid}, Ptr{Cvoid}), prob.h, A))
    return nothing
end

MHDFlows.stepforward!(prob::MI355XProblem) =
    check(ccall((:mhd_step, libmhd), Int32, (Ptr{Cvoid}, Int32), prob.h, 1))

"The diffusive time-step bound of TimeIntegrator! (src/integrator.jl:60-72)."
function diffusion_dt(prob::MI355XProblem, CFL_Coef)
    p, g = prob.params, prob.grid
    vi = prob.flag.b ? max(p.ν, p.η) : p.ν
    nv = prob.flag.b ? max(p.nν, p.nη) : p.nν
    dl = min(g.Lx / g.nx, g.Ly / g.ny, g.Lz / g.nz)
    return nv > 1 ? CFL_Coef * dl^nv / vi : CFL_Coef * dl^2 / vi
end

function MHDFlows.TimeIntegrator!(prob::MI355XProblem, t₀::Number, N₀::Int; usr_dt=0.0, CFL_Coef=0.25,
                                  CFL_function=MHDFlows.nothingfunction, diags=[], save=false,
                                  save_loc="", filename="", file_number=0, dump_dt=0, kwargs...)
    file_path_and_name = ""
    if save                                                      # integrator.jl:44-51
        if length(save_loc) == 0 || length(filename) == 0 || dump_dt == 0
            error("Save Function Turned ON but save_loc/filename/dump_dt is not declared!\n")
        end
        file_path_and_name = save_loc * filename
        savefile(prob, file_number; file_path_and_name=file_path_and_name)
        file_number += 1
    end
    user_cfl = CFL_function != MHDFlows.nothingfunction
    if user_cfl && usr_dt > 0.0                                  # CFL_Init, :200-206
        error("User define both CFL_function and usr_dt")
    end
    hooks = [f for f in prob.usr_func if f != MHDFlows.nothingfunction]
    vp = prob.flag.vp
    if isempty(diags) && isempty(hooks) && !save && !user_cfl && !vp
        # the whole loop on the device: dt from getCFL! there, no host synchronisation
        steps = Ref{Int64}(0)
        check(ccall((:mhd_integrate, libmhd), Int32,
                    (Ptr{Cvoid}, Float64, Int64, Float64, Float64, Ref{Int64}),
                    prob.h, Float64(t₀), N₀, Float64(usr_dt), Float64(CFL_Coef), steps))
        return nothing
    end
    # per-step loop with host hooks (integrator.jl:104-146)
    t_diff = diffusion_dt(prob, CFL_Coef)
    t_next_save = prob.clock.t + dump_dt
    prob.clock.step = 0
    usr_dt != 0.0 && (prob.clock.dt = usr_dt)
    if vp                                                        # :84-87
        MHDFlows.DivVCorrection!(prob)
        prob.flag.b && MHDFlows.DivBCorrection!(prob)
    end
    while N₀ >= prob.clock.step && t₀ >= prob.clock.t
        if user_cfl
            CFL_function(prob, t_diff; Coef=CFL_Coef)           # sets prob.clock.dt
        elseif usr_dt == 0.0
            MHDFlows.getCFL!(prob, t_diff; Coef=CFL_Coef)
        end
        MHDFlows.stepforward!(prob)
        for d in diags                                           # increment!(diags)
            FourierFlows.increment!(d)
        end
        if vp                                                    # :118-121
            MHDFlows.DivVCorrection!(prob)
            prob.flag.b && MHDFlows.DivBCorrection!(prob)
        end
        for foo! in hooks                                        # usr_func
            foo!(prob)
        end
        if save && prob.clock.t >= t_next_save                  # :136-141
            MHDFlows.ProbDiagnostic(prob)
            savefile(prob, file_number; file_path_and_name=file_path_and_name)
            t_next_save += dump_dt
            file_number += 1
        end
    end
    return nothing
end

function MHDFlows.getCFL!(prob::MI355XProblem, t_diff; Coef=0.3)
    dt = Ref{Float64}(0.0)
    check(ccall((:mhd_cfl_dt, libmhd), Int32, (Ptr{Cvoid}, Float64, Ref{Float64}), prob.h, Coef, dt))
    dt[] = min(dt[], t_diff)                                     # dt = minimum([Coef*dl/vmax, t_diff])
    prob.clock.dt = dt[]
    return dt[]
end

# Checkpoints: the library's container (libhdf5 is not linked); datasets keep the reference's
# names (i/j/k_velocity, i/j/k_mag_field, time).  exact=true adds the spectral state and clock.
function savefile(prob::MI355XProblem, file_number; file_path_and_name="", exact=false)
    name = file_path_and_name * "_t_" * lpad(string(file_number), 4, '0') * ".mhdf"
    check(ccall((:mhd_savefile, libmhd), Int32, (Ptr{Cvoid}, Cstring, Int32), prob.h, name, Int32(exact)))
    return name
end

function Restart!(prob::MI355XProblem, file_path_and_name; exact=false)
    check(ccall((:mhd_restart, libmhd), Int32, (Ptr{Cvoid}, Cstring, Int32), prob.h, file_path_and_name,
                Int32(exact)))
    return nothing
end

# Dealiased steps (default on): every step of a call but the last updates kept modes only;
# the state after the call is the same as with every step computed in full.
set_step_pruning!(prob::MI355XProblem, enable::Bool) =
    check(ccall((:mhd_set_step_pruning, libmhd), Int32, (Ptr{Cvoid}, Int32), prob.h, Int32(enable)))

function autotune!(prob::MI355XProblem; reps=5)
    buf = Vector{UInt8}(undef, 4096)
    check(ccall((:mhd_autotune, libmhd), Int32, (Ptr{Cvoid}, Int32, Ptr{UInt8}, Int32), prob.h, reps, buf, 4096))
    return unsafe_string(pointer(buf))
end

# Forcing modules (the calcF! hook, src/pgen.jl:231-234), evaluated inside the device step
realptrs(prob::MI355XProblem{T}, fs) where T =
    [f === nothing ? nothing : convert(Array{T,3}, f) for f in fs]

function SetUpN97!(prob::MI355XProblem{T}; F0=1, kf=2) where T
    g = prob.grid
    X = reshape(g.x, :, 1, 1); Y = reshape(g.y, 1, :, 1); Z = reshape(g.z, 1, 1, :)
    k, F = T(kf), T(F0)
    fx = F .* sin.(k .* X) .* cos.(k .* Y) .* cos.(k .* Z)
    fy = F .* -cos.(k .* X) .* sin.(k .* Y) .* cos.(k .* Z)
    arrays = realptrs(prob, (fx, fy, nothing))
    ptrs = Ptr{Cvoid}[a === nothing ? C_NULL : pointer(a) for a in arrays]
    GC.@preserve arrays check(ccall((:mhd_set_forcing_field, libmhd), Int32, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}),
                                    prob.h, ptrs))
end

function SetUpND!(prob::MI355XProblem, P, fx, fy, fz)
    arrays = realptrs(prob, (fx, fy, fz))
    ptrs = Ptr{Cvoid}[pointer(a) for a in arrays]
    GC.@preserve arrays check(ccall((:mhd_set_forcing_nd, libmhd), Int32, (Ptr{Cvoid}, Float64, Ptr{Ptr{Cvoid}}),
                                    prob.h, Float64(P), ptrs))
end

# `seed` keys the counter-based draws that replace the reference's unseeded rand()
SetUpFk(prob::MI355XProblem; kf=2, P=1, σ²=1, seed=0, b=1.0, A=1.0) =
    check(ccall((:mhd_set_forcing_a99, libmhd), Int32,
                (Ptr{Cvoid}, UInt64, Float64, Float64, Float64, Float64, Float64),
                prob.h, UInt64(seed), Float64(kf), Float64(P), Float64(σ²), Float64(b), Float64(A)))

# Analysis of prob.vars on the device (src/utils/MHDAnalysis.jl)
"Shell spectra of every vars field: a (krmax, nf + 1) matrix, the last column the mode count."
function spectra(prob::MI355XProblem)
    nb = Ref{Int32}(0)
    check(ccall((:mhd_spectra, libmhd), Int32, (Ptr{Cvoid}, Ptr{Float64}, Int32, Ref{Int32}), prob.h, C_NULL, 0, nb))
    out = zeros(Float64, nb[], prob.nf + 1)
    check(ccall((:mhd_spectra, libmhd), Int32, (Ptr{Cvoid}, Ptr{Float64}, Int32, Ref{Int32}), prob.h, out, nb[], nb))
    return out
end

"spectralline of a vars field (MHDAnalysis.jl:236-255): (Pk, kr) with kr the 1-based shell label."
function MHDFlows.spectralline(prob::MI355XProblem{T}, field::Symbol=:ux) where T
    s = spectra(prob)
    f = findfirst(==(field), FIELDS)
    kr = T[s[r, end] > 0 ? r : 0 for r in 1:size(s, 1)]
    return s[:, f], kr
end

"ScaleDecomposition of consecutive vars fields (MHDAnalysis.jl:17-82): k1 <= |k| <= k2 kept."
function MHDFlows.ScaleDecomposition(prob::MI355XProblem{T}, fields::Tuple; kf=[1, 5]) where T
    f0 = findfirst(==(fields[1]), FIELDS) - 1
    g = prob.grid
    out = Array{T,4}(undef, g.nx, g.ny, g.nz, length(fields))
    check(ccall((:mhd_scale_decomposition, libmhd), Int32, (Ptr{Cvoid}, Int32, Int32, Float64, Float64, Ptr{Cvoid}),
                prob.h, f0, length(fields), minimum(kf), maximum(kf), out))
    return Tuple(out[:, :, :, c] for c in 1:length(fields))
end

# Vector calculus of host fields on the device (src/utils/VectorCalculus.jl, MHDAnalysis.jl):
# the problem's grid and precision; op codes of mhd_field_op
function fieldop(prob::MI355XProblem{T}, op::Integer, fields, nout::Integer) where T
    g = prob.grid
    ins = [convert(Array{T,3}, f) for f in fields]
    outs = [Array{T,3}(undef, g.nx, g.ny, g.nz) for _ in 1:nout]
    pin = Ptr{Cvoid}[pointer(a) for a in ins]
    pout = Ptr{Cvoid}[pointer(a) for a in outs]
    GC.@preserve ins outs check(ccall((:mhd_field_op, libmhd), Int32, (Ptr{Cvoid}, Int32, Ptr{Ptr{Cvoid}},
                                      Ptr{Ptr{Cvoid}}), prob.h, op, pin, pout))
    return outs
end
MHDFlows.Curl(prob::MI355XProblem, B1, B2, B3) = Tuple(fieldop(prob, 7, (B1, B2, B3), 3))
MHDFlows.Div(prob::MI355XProblem, B1, B2, B3) = fieldop(prob, 8, (B1, B2, B3), 1)[1]
MHDFlows.∂i(prob::MI355XProblem, B1, direction) =
    fieldop(prob, direction == "x" ? 3 : direction == "y" ? 4 : direction == "z" ? 5 :
                  error("Wrong driection declared"), (B1,), 1)[1]
MHDFlows.LaplaceSolver(prob::MI355XProblem, B) = fieldop(prob, 6, (B,), 1)[1]
MHDFlows.VectorPotential(prob::MI355XProblem, B1, B2, B3) = Tuple(fieldop(prob, 9, (B1, B2, B3), 3))

# DivVCorrection! / DivBCorrection! (src/Solver/VPSolver.jl:60-139, exported by MHDFlows)
MHDFlows.DivVCorrection!(prob::MI355XProblem) =
    check(ccall((:mhd_div_correction, libmhd), Int32, (Ptr{Cvoid}, Int32), prob.h, 0))
MHDFlows.DivBCorrection!(prob::MI355XProblem) =
    check(ccall((:mhd_div_correction, libmhd), Int32, (Ptr{Cvoid}, Int32), prob.h, 1))

"""Velocity structure functions of prob.vars (src/utils/TurbStatTool.jl): mode "1D" = SF₂1D
(vector of 2R bins), "Global" = SFr_Global against the mean magnetic field ((R+1) x (R+1)
matrix SFVr[rpar, rperp]).  Returns (values, counts); empty bins are NaN."""
function SFr(prob::MI355XProblem, mode::String="1D")
    m = mode == "1D" ? 0 : mode == "Global" ? 1 : error("mode must be \"1D\" or \"Global\"")
    nb = Ref{Int32}(0); nrow = Ref{Int32}(0)
    f = (:mhd_structure_function, libmhd)
    check(ccall(f, Int32, (Ptr{Cvoid}, Int32, Ptr{Float64}, Int32, Ref{Int32}, Ref{Int32}),
                prob.h, m, C_NULL, 0, nb, nrow))
    out = Vector{Float64}(undef, 2 * nb[])
    check(ccall(f, Int32, (Ptr{Cvoid}, Int32, Ptr{Float64}, Int32, Ref{Int32}, Ref{Int32}),
                prob.h, m, out, length(out), nb, nrow))
    vals, cnt = out[1:nb[]], out[nb[]+1:end]
    m == 1 && return reshape(vals, nrow[], nrow[]), reshape(cnt, nrow[], nrow[])
    return vals, cnt
end

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
