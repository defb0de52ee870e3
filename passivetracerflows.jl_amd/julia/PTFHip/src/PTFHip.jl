"""
    PTFHip

A `ROCm` device for PassiveTracerFlows.jl, backed by libptf_hip (include/ptf.h)
through `ccall`.  With it the reference's own scripts run on an MI355X by changing
the device only:

    using PassiveTracerFlows, PTFHip
    dev = ROCm()                                      # instead of CPU() / GPU()
    prob = TracerAdvectionDiffusion.Problem(dev, TwoDAdvectingFlow(; u, v); nx=256, Lx=2π,
                                            κ=2e-3, dt=0.01, stepper="RK4")
    sol, clock, vars, params, grid = prob.sol, prob.clock, prob.vars, prob.params, prob.grid
    TracerAdvectionDiffusion.set_c!(prob, c0)
    stepforward!(prob, 2000)
    TracerAdvectionDiffusion.updatevars!(prob)        # prob.vars.c

How the reference's code reaches the library (every method below dispatches on a type
of this package, so nothing of PassiveTracerFlows or FourierFlows is overwritten):

* `TwoDGrid(dev::ROCm; ...)` (and 1D/3D) builds FourierFlows' own grid type (so the
  reference's `grid::TwoDGrid` methods apply) with `device = dev` and
  `rfftplan = PtfRFFTPlan`, a libptf_hip plan; `mul!`/`ldiv!` on it run on the GPU.
  Coordinates and wavenumbers stay host arrays (`device_array(::ROCm) = Array`).
* `zeros(::ROCm, T, dims)` gives host arrays for real types and `PtfSol` (a host
  mirror of a device spectral state) for complex types, so the `sol` that
  `FourierFlows.Problem` allocates is a `PtfSol`.
* `FourierFlows.TimeStepper(stepper, eqn, dt, ::ROCm)` gives a `PtfTimeStepper`;
  `stepforward!(sol::PtfSol, clock, ::PtfTimeStepper, eqn, vars, params, grid)` runs the
  step on the GPU (one `ccall`), with the flow at `clock.t`, then advances the clock as
  FourierFlows does.  The first call creates the libptf_hip problem from `params`,
  `grid`, the stepper and `clock.dt` (src/traceradvectiondiffusion.jl:143-216, 357-554).
* `set_c!` / `updatevars!` with a `PtfSol` state use ptf_set_c / ptf_updatevars once a
  device problem exists; before that, the reference's versions run through the plan.

`sol` stays device resident: reading `prob.sol` downloads the state once after each
step (lazily); writing it uploads before the next step.
"""
module PTFHip

using LinearAlgebra
using FourierFlows
using PassiveTracerFlows
import PassiveTracerFlows.TracerAdvectionDiffusion as TAD
using PassiveTracerFlows.TracerAdvectionDiffusion: AbstractSteadyFlowParams, AbstractTimeVaryingFlowParams

import LinearAlgebra: mul!, ldiv!

export ROCm, PtfArray, PtfSol

const libptf = get(ENV, "PTF_LIB", joinpath(@__DIR__, "..", "..", "..", "lib", "libptf_hip.so"))

# ---- C ABI mirror (include/ptf.h) ------------------------------------------------
const PTF_FP64, PTF_MIXED = Int32(0), Int32(1)
const PTF_RK4, PTF_ETDRK4, PTF_FILTERED_RK4, PTF_FILTERED_ETDRK4 = Int32.((0, 1, 2, 3))
const STEPPERS = Dict("RK4" => PTF_RK4, "ETDRK4" => PTF_ETDRK4, "FilteredRK4" => PTF_FILTERED_RK4,
                      "FilteredETDRK4" => PTF_FILTERED_ETDRK4)
const PTF_FLOW_STEADY, PTF_FLOW_PER_STEP, PTF_FLOW_SPECTRAL, PTF_FLOW_ANALYTIC = Int32.((0, 1, 2, 3))

"`ptf_desc` of include/ptf.h (field order, types and C layout; checked by tests/test_julia_shim.py)."
struct PtfDesc
  ndim::Int32
  precision::Int32
  n::NTuple{3,Int64}
  L::NTuple{3,Float64}
  kappa::Float64
  eta::Float64
  iota::Float64
  kappah::Float64
  nkappah::Int32
  stepper::Int32
  dt::Float64
  flow_kind::Int32
  filter_order::Int32
  filter_innerK::Float64
  filter_outerK::Float64
  filter_tol::Float64
end

"One separable term of a built-in analytic flow (`ptf_flow_term` of include/ptf.h)."
struct PtfFlowTerm
  component::Int32
  fx::Int32
  fy::Int32
  fz::Int32
  kx::Float64
  ky::Float64
  kz::Float64
  px::Float64
  py::Float64
  pz::Float64
  amp::Float64
  time_fn::Int32
  reserved::Int32
  omega::Float64
  tphase::Float64
end

lasterror() = unsafe_string(ccall((:ptf_last_error, libptf), Cstring, ()))
check(rc) = rc == 0 || error("libptf_hip error $rc: $(lasterror())")

# ---- the device ---------------------------------------------------------------------
"One MI355X (HIP device `ordinal`)."
struct ROCm <: FourierFlows.Device
  ordinal::Int
  function ROCm(ordinal::Integer=0)
    check(ccall((:ptf_device_init, libptf), Cint, (Int32,), ordinal))
    return new(ordinal)
  end
end

# grid metadata (coordinates, wavenumbers, L, params) are host arrays; the library keeps
# its own device copies of what the step needs
FourierFlows.device_array(::ROCm) = Array
FourierFlows.device_array(::ROCm, T, dim) = Array{T, dim}

"Device buffer owned by libptf_hip (freed by a finalizer); for `mul!`/`ldiv!` on device data."
mutable struct PtfArray{T,N} <: AbstractArray{T,N}
  ptr::Ptr{Cvoid}
  dims::NTuple{N,Int}
  function PtfArray{T}(::UndefInitializer, dims::NTuple{N,Int}) where {T,N}
    p = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:ptf_buffer_alloc, libptf), Cint, (UInt64, Ref{Ptr{Cvoid}}), prod(dims) * sizeof(T), p))
    a = new{T,N}(p[], dims)
    finalizer(x -> ccall((:ptf_buffer_free, libptf), Cint, (Ptr{Cvoid},), x.ptr), a)
    return a
  end
end
PtfArray(a::Array{T,N}) where {T,N} = copyto!(PtfArray{T}(undef, size(a)), a)
Base.size(a::PtfArray) = a.dims
Base.getindex(a::PtfArray, i::Int...) = Array(a)[i...]            # slow: whole download
function Base.copyto!(dst::PtfArray{T}, src::Array{T}) where T
  length(dst) == length(src) || throw(DimensionMismatch("size mismatch"))
  GC.@preserve src check(ccall((:ptf_buffer_upload, libptf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, UInt64),
                               dst.ptr, pointer(src), sizeof(src)))
  return dst
end
function Base.Array(a::PtfArray{T,N}) where {T,N}
  out = Array{T,N}(undef, a.dims)
  check(ccall((:ptf_buffer_download, libptf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, UInt64),
              pointer(out), a.ptr, sizeof(out)))
  return out
end

# ---- the plan: grid.rfftplan with mul!/ldiv! -------------------------------------------
"libptf_hip real-to-complex plan over all grid dimensions (FFTW's rfft plan on ROCm)."
mutable struct PtfRFFTPlan{T}
  h::Ptr{Cvoid}
  ndim::Int
  n::NTuple{3,Int}
  function PtfRFFTPlan{T}(ndim::Integer, n::NTuple{3,Int}) where T
    h = Ref{Ptr{Cvoid}}(C_NULL)
    prec = T == Float64 ? PTF_FP64 : PTF_MIXED
    check(ccall((:ptf_plan_create, libptf), Cint, (Int32, Ref{NTuple{3,Int64}}, Int32, Ref{Ptr{Cvoid}}),
                ndim, Int64.(n), prec, h))
    p = new{T}(h[], ndim, n)
    finalizer(x -> ccall((:ptf_plan_destroy, libptf), Cint, (Ptr{Cvoid},), x.h), p)
    return p
  end
end
physdims(p::PtfRFFTPlan) = p.n[1:p.ndim]
specdims(p::PtfRFFTPlan) = (p.n[1] ÷ 2 + 1, p.n[2:p.ndim]...)

function checkdims(a, dims, what)
  size(a) == dims || throw(DimensionMismatch("$what has size $(size(a)), the plan needs $dims"))
end

"mul!(out, plan, a): unnormalised R2C on the GPU (S:712-778, S:847)."
function mul!(out::PtfArray{Complex{T}}, p::PtfRFFTPlan{T}, a::PtfArray{T}) where T
  checkdims(a, physdims(p), "input"); checkdims(out, specdims(p), "output")
  check(ccall((:ptf_rfft, libptf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}), p.h, a.ptr, out.ptr))
  return out
end
"ldiv!(out, plan, a): C2R scaled by 1/N on the GPU; the input is preserved (S:698, S:818)."
function ldiv!(out::PtfArray{T}, p::PtfRFFTPlan{T}, a::PtfArray{Complex{T}}) where T
  checkdims(a, specdims(p), "input"); checkdims(out, physdims(p), "output")
  check(ccall((:ptf_irfft, libptf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}), p.h, a.ptr, out.ptr))
  return out
end
# host arrays (the grid's and vars' own): staged through device buffers
function mul!(out::AbstractArray{<:Complex}, p::PtfRFFTPlan{T}, a::AbstractArray{<:Real}) where T
  d = PtfArray{Complex{T}}(undef, specdims(p))
  mul!(d, p, PtfArray(Array{T}(a)))
  copyto!(out, Array(d))
  return out
end
function ldiv!(out::AbstractArray{<:Real}, p::PtfRFFTPlan{T}, a::AbstractArray{<:Complex}) where T
  d = PtfArray{T}(undef, physdims(p))
  ldiv!(d, p, PtfArray(Array{Complex{T}}(a)))
  copyto!(out, Array(d))
  return out
end

# ---- grids: FourierFlows' own types with the ROCm device and plan --------------------
# Rebuild a grid of FourierFlows' type G with `device` and `rfftplan` replaced (by field
# name, so the field order of the FourierFlows release does not matter).
function with_device_and_plan(g::G, dev, plan) where G
  vals = map(f -> f === :device ? dev : (f === :rfftplan ? plan : getfield(g, f)), fieldnames(G))
  params = collect(G.parameters)
  i = findfirst(p -> p === fieldtype(G, :rfftplan), params)
  i === nothing && error("unexpected grid type $G: rfftplan is not a type parameter")
  params[i] = typeof(plan)
  return G.name.wrapper{params...}(vals...)
end

function FourierFlows.OneDGrid(dev::ROCm; nx, Lx, T=Float64, kw...)
  g = FourierFlows.OneDGrid(FourierFlows.CPU(); nx, Lx, T, kw...)
  return with_device_and_plan(g, dev, PtfRFFTPlan{T}(1, (nx, 1, 1)))
end
function FourierFlows.TwoDGrid(dev::ROCm; nx, Lx, ny=nx, Ly=Lx, T=Float64, kw...)
  g = FourierFlows.TwoDGrid(FourierFlows.CPU(); nx, Lx, ny, Ly, T, kw...)
  return with_device_and_plan(g, dev, PtfRFFTPlan{T}(2, (nx, ny, 1)))
end
function FourierFlows.ThreeDGrid(dev::ROCm; nx, Lx, ny=nx, Ly=Lx, nz=nx, Lz=Lx, T=Float64, kw...)
  g = FourierFlows.ThreeDGrid(FourierFlows.CPU(); nx, Lx, ny, Ly, nz, Lz, T, kw...)
  return with_device_and_plan(g, dev, PtfRFFTPlan{T}(3, (nx, ny, nz)))
end

# ---- the spectral state ------------------------------------------------------------
"""
    PtfSol{T,N}

Host mirror of a spectral state that lives in a libptf_hip problem.  Reads refresh the
mirror once after the device has stepped (`stale`); writes mark it `dirty`, and the next
step uploads it first.  Until a device problem is attached it is a plain host array.
"""
mutable struct PtfSol{T,N} <: AbstractArray{T,N}
  host::Array{T,N}
  dev::Any          # the DeviceProblem whose state this is, or nothing
  stale::Bool
  dirty::Bool
end
PtfSol(a::Array{T,N}) where {T,N} = PtfSol{T,N}(a, nothing, false, false)

function refresh!(a::PtfSol)
  if a.stale && a.dev !== nothing
    check(ccall((:ptf_get_sol, libptf), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}), a.dev.h, a.dev.buf))
    copyto!(a.host, a.dev.buf)
    a.stale = false
  end
  return a.host
end
Base.size(a::PtfSol) = size(a.host)
Base.IndexStyle(::Type{<:PtfSol}) = IndexLinear()
Base.getindex(a::PtfSol, i::Int) = refresh!(a)[i]
function Base.setindex!(a::PtfSol, v, i::Int)
  refresh!(a)[i] = v
  a.dirty = true
  return v
end
Base.similar(a::PtfSol, ::Type{S}, dims::Dims) where S = PtfSol(similar(a.host, S, dims))
Base.Array(a::PtfSol) = copy(refresh!(a))
Base.unsafe_convert(::Type{Ptr{T}}, a::PtfSol{T}) where T = Base.unsafe_convert(Ptr{T}, refresh!(a))
Base.BroadcastStyle(::Type{<:PtfSol}) = Broadcast.ArrayStyle{PtfSol}()
function Base.similar(bc::Broadcast.Broadcasted{Broadcast.ArrayStyle{PtfSol}}, ::Type{S}) where S
  return PtfSol(similar(Array{S}, axes(bc)))
end
function Base.copyto!(dst::PtfSol, bc::Broadcast.Broadcasted)
  copyto!(refresh!(dst), bc)
  dst.dirty = true
  return dst
end
function Base.copyto!(dst::PtfSol, src::AbstractArray)
  copyto!(refresh!(dst), src)
  dst.dirty = true
  return dst
end

Base.zeros(::ROCm, ::Type{T}, dims) where {T<:Complex} = PtfSol(zeros(T, dims))
Base.zeros(::ROCm, ::Type{T}, dims) where {T} = zeros(T, dims)

# ---- the time stepper and the device problem --------------------------------------------
"FourierFlows time stepper on ROCm: the stage algebra runs inside libptf_hip."
mutable struct PtfTimeStepper{A} <: FourierFlows.AbstractTimeStepper{A}
  stepper::String
  filter::NamedTuple
end

function FourierFlows.TimeStepper(stepper, eqn, dt, dev::ROCm; order=0, innerK=0.65, outerK=1.0,
                                  tol=1e-15)
  name = replace(string(stepper), "TimeStepper" => "")
  haskey(STEPPERS, name) ||
    throw(ArgumentError("stepper $stepper is not implemented on ROCm (RK4, ETDRK4, FilteredRK4, FilteredETDRK4 are)"))
  return PtfTimeStepper{PtfSol{eqn.T, length(eqn.dims)}}(name, (; order, innerK, outerK, tol))
end

mutable struct DeviceProblem
  h::Ptr{Cvoid}
  buf::Array{ComplexF64}       # FP64 staging of sol in the reference layout
  steady::Bool
  grid::Any
end

field(p, name, default) = hasproperty(p, name) ? getproperty(p, name) : default

function flow_kind(params)
  params isa AbstractSteadyFlowParams && return PTF_FLOW_STEADY
  params isa AbstractTimeVaryingFlowParams && return PTF_FLOW_PER_STEP
  throw(ArgumentError("$(typeof(params)) is not on the ROCm path (steady or time-varying flows are)"))
end

"The libptf_hip problem behind `sol`, created on first use (S:143-216 as a ptf_desc)."
function device_problem!(sol::PtfSol, ts::PtfTimeStepper, dt, params, grid)
  sol.dev === nothing || return sol.dev
  nd = ndims(sol)
  n = nd == 1 ? (grid.nx, 1, 1) : nd == 2 ? (grid.nx, grid.ny, 1) : (grid.nx, grid.ny, grid.nz)
  L = nd == 1 ? (grid.Lx, 1.0, 1.0) : nd == 2 ? (grid.Lx, grid.Ly, 1.0) : (grid.Lx, grid.Ly, grid.Lz)
  κ = params.κ
  f = ts.filter
  d = PtfDesc(nd, eltype(grid) == Float64 ? PTF_FP64 : PTF_MIXED, Int64.(n), Float64.(L),
              κ, field(params, :η, κ), field(params, :ι, κ), params.κh, params.nκh,
              STEPPERS[ts.stepper], dt, flow_kind(params), f.order, f.innerK, f.outerK, f.tol)
  h = Ref{Ptr{Cvoid}}(C_NULL)
  check(ccall((:ptf_problem_create, libptf), Cint, (Ref{PtfDesc}, Ref{Ptr{Cvoid}}), d, h))
  dp = DeviceProblem(h[], zeros(ComplexF64, size(sol)), params isa AbstractSteadyFlowParams, grid)
  finalizer(x -> ccall((:ptf_problem_destroy, libptf), Cint, (Ptr{Cvoid},), x.h), dp)
  if dp.steady
    upload_flow!(dp, [Float64.(field(params, c, nothing)) for c in (:u, :v, :w)[1:nd]])
  end
  sol.dev = dp
  sol.dirty = true            # the host state is the initial state
  return dp
end

function upload_flow!(dp::DeviceProblem, fields)
  ptrs = Ptr{Float64}[pointer(a) for a in fields]
  while length(ptrs) < 3
    push!(ptrs, Ptr{Float64}(C_NULL))
  end
  GC.@preserve fields check(ccall((:ptf_set_flow, libptf), Cint,
                                  (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                                  dp.h, ptrs[1], ptrs[2], ptrs[3]))
end

"A time-varying user flow evaluated on the host at `t` (calcN! uses clock.t: S:701, 718, 737)."
function upload_flow_at!(dp::DeviceProblem, params, grid, t)
  pts = gridpoints(grid)
  pts = pts isa Tuple ? pts : (pts,)
  fs = [getproperty(params, c) for c in (:u, :v, :w)[1:length(pts)]]
  upload_flow!(dp, [Float64.(f.(pts..., t)) .+ zeros(size(pts[1])) for f in fs])
end

function push_state!(sol::PtfSol)
  if sol.dirty
    copyto!(sol.dev.buf, sol.host)
    check(ccall((:ptf_set_sol, libptf), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}), sol.dev.h, sol.dev.buf))
    sol.dirty = false
  end
end

"""
    stepforward!(sol::PtfSol, clock, ts::PtfTimeStepper, eqn, vars, params, grid)

One step of FourierFlows' `stepforward!` on the GPU: RK4/ETDRK4 (filtered or not) with the
fused calcN! (S:695-786), the flow frozen at `clock.t`, then `clock.t += dt; clock.step += 1`.
"""
function FourierFlows.stepforward!(sol::PtfSol, clock, ts::PtfTimeStepper, eqn, vars, params, grid)
  dp = device_problem!(sol, ts, clock.dt, params, grid)
  push_state!(sol)
  check(ccall((:ptf_set_clock, libptf), Cint, (Ptr{Cvoid}, Float64, Int64), dp.h, clock.t, clock.step))
  dp.steady || upload_flow_at!(dp, params, grid, clock.t)
  check(ccall((:ptf_step, libptf), Cint, (Ptr{Cvoid}, Int64), dp.h, 1))
  clock.t += clock.dt
  clock.step += 1
  sol.stale = true
  return nothing
end

"""
    step_n!(prob, nsteps)

`stepforward!(prob, nsteps)` in one `ccall` for a steady flow (the library replays captured
graphs of 8 steps); a time-varying flow is stepped one step at a time.
"""
function step_n!(prob, nsteps::Integer)
  sol, clock = prob.sol, prob.clock
  sol isa PtfSol || return stepforward!(prob, nsteps)
  dp = device_problem!(sol, prob.timestepper, clock.dt, prob.params, prob.grid)
  if !dp.steady
    for _ in 1:nsteps
      stepforward!(prob)
    end
    return nothing
  end
  push_state!(sol)
  check(ccall((:ptf_set_clock, libptf), Cint, (Ptr{Cvoid}, Float64, Int64), dp.h, clock.t, clock.step))
  check(ccall((:ptf_step, libptf), Cint, (Ptr{Cvoid}, Int64), dp.h, nsteps))
  t = Ref{Float64}(0); s = Ref{Int64}(0)
  check(ccall((:ptf_get_clock, libptf), Cint, (Ptr{Cvoid}, Ref{Float64}, Ref{Int64}), dp.h, t, s))
  clock.t, clock.step = t[], s[]
  sol.stale = true
  return nothing
end

# ---- set_c! / updatevars! (S:815-872) with a device state ---------------------------------
const FlowParams = Union{AbstractSteadyFlowParams, AbstractTimeVaryingFlowParams}

function TAD.updatevars!(params::FlowParams, vars, grid, sol::PtfSol)
  dp = sol.dev
  if dp === nothing
    @. vars.ch = sol
    ldiv!(vars.c, grid.rfftplan, deepcopy(vars.ch))
    return nothing
  end
  push_state!(sol)
  c = zeros(Float64, size(vars.c))
  check(ccall((:ptf_updatevars, libptf), Cint, (Ptr{Cvoid}, Ptr{Float64}), dp.h, c))
  copyto!(vars.c, c)
  vars.ch .= refresh!(sol)
  return nothing
end

function TAD.set_c!(sol::PtfSol, params::FlowParams, vars, grid, c)
  dp = sol.dev
  if dp === nothing
    mul!(sol, grid.rfftplan, Array(c))
  else
    cc = Float64.(Array(c))
    GC.@preserve cc check(ccall((:ptf_set_c, libptf), Cint, (Ptr{Cvoid}, Ptr{Float64}), dp.h, pointer(cc)))
    sol.stale = true
    sol.dirty = false
  end
  TAD.updatevars!(params, vars, grid, sol)
  return nothing
end

# ---- GPU-resident diagnostics (ptf_step_record) -------------------------------------
"""
    step_record!(prob, nsteps, freq)

Step `nsteps` on the GPU and record t, step, mass and variance every `freq` steps there
(FourierFlows Diagnostic schedule); one copy to the host at the end.
"""
function step_record!(prob, nsteps::Integer, freq::Integer)
  sol, clock = prob.sol, prob.clock
  dp = device_problem!(sol, prob.timestepper, clock.dt, prob.params, prob.grid)
  dp.steady || throw(ArgumentError("step_record! needs a steady flow"))
  push_state!(sol)
  check(ccall((:ptf_set_clock, libptf), Cint, (Ptr{Cvoid}, Float64, Int64), dp.h, clock.t, clock.step))
  nrec = max(div(clock.step + nsteps, freq) - div(clock.step, freq), 1)
  t = zeros(nrec); st = zeros(Int64, nrec); m = zeros(nrec); v = zeros(nrec)
  n = Ref{Int64}(0)
  check(ccall((:ptf_step_record, libptf), Cint,
              (Ptr{Cvoid}, Int64, Int64, Int64, Ref{Int64}, Ptr{Float64}, Ptr{Int64}, Ptr{Float64},
               Ptr{Float64}, Ptr{Float64}),
              dp.h, nsteps, freq, nrec, n, t, st, m, v, C_NULL))
  tt = Ref{Float64}(0); ss = Ref{Int64}(0)
  check(ccall((:ptf_get_clock, libptf), Cint, (Ptr{Cvoid}, Ref{Float64}, Ref{Int64}), dp.h, tt, ss))
  clock.t, clock.step = tt[], ss[]
  sol.stale = true
  k = n[]
  return (t=t[1:k], steps=st[1:k], mass=m[1:k], variance=v[1:k])
end

end # module
