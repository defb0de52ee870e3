"""
    PTFHip

A `ROCm` device for PassiveTracerFlows.jl's `TracerAdvectionDiffusion` time step,
backed by libptf_hip (include/ptf.h) through `ccall`.

    using PassiveTracerFlows, PTFHip
    dev  = PTFHip.ROCm()
    prob = PTFHip.Problem(dev, TwoDAdvectingFlow(; u, v); nx=256, Lx=2π, κ=2e-3, dt=0.01)
    set_c!(prob, c0); stepforward!(prob, 2000); updatevars!(prob); prob.vars.c

`stepforward!` makes one `ccall` per call (all steps run on the GPU) and then
mirrors the library clock into `prob.clock` (t += dt, step += 1 per step, after
the step, as FourierFlows does).  A time-varying user flow is evaluated on the
host at `clock.t` before every step and uploaded (calcN! evaluates the flow at
`clock.t`, src/traceradvectiondiffusion.jl:701, 718, 737).
"""
module PTFHip

using PassiveTracerFlows
using PassiveTracerFlows.TracerAdvectionDiffusion: OneDAdvectingFlow, TwoDAdvectingFlow,
      ThreeDAdvectingFlow, noflow
import PassiveTracerFlows.TracerAdvectionDiffusion: set_c!, updatevars!
import FourierFlows: stepforward!
import LinearAlgebra: mul!, ldiv!

const libptf = get(ENV, "PTF_LIB", joinpath(@__DIR__, "..", "..", "..", "lib", "libptf_hip.so"))

# ---- C ABI mirror (include/ptf.h) ------------------------------------------------
const PTF_FP64, PTF_MIXED = Int32(0), Int32(1)
const PTF_RK4, PTF_ETDRK4, PTF_FILTERED_RK4, PTF_FILTERED_ETDRK4 = Int32.((0, 1, 2, 3))
const STEPPERS = Dict("RK4" => PTF_RK4, "ETDRK4" => PTF_ETDRK4, "FilteredRK4" => PTF_FILTERED_RK4,
                      "FilteredETDRK4" => PTF_FILTERED_ETDRK4)
const PTF_FLOW_STEADY, PTF_FLOW_PER_STEP, PTF_FLOW_SPECTRAL = Int32(0), Int32(1), Int32(2)

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
  filter_order::Int32          # 0: FourierFlows' makefilter defaults
  filter_innerK::Float64
  filter_outerK::Float64
  filter_tol::Float64
end

"One separable term of a built-in analytic flow (ptf_flow_term in include/ptf.h)."
struct PtfFlowTerm
  component::Int32            # 0 = u, 1 = v, 2 = w
  fx::Int32; fy::Int32; fz::Int32   # 0 = 1, 1 = sin, 2 = cos
  kx::Float64; ky::Float64; kz::Float64
  px::Float64; py::Float64; pz::Float64
  amp::Float64
  time_fn::Int32
  reserved::Int32
  omega::Float64
  tphase::Float64
end
const PTF_FLOW_ANALYTIC = Int32(3)

lasterror() = unsafe_string(ccall((:ptf_last_error, libptf), Cstring, ()))
check(rc) = rc == 0 || error("libptf_hip error $rc: $(lasterror())")

# ---- device and plans (FourierFlows Device, grid.rfftplan with mul!/ldiv!) -----------
struct ROCm <: FourierFlows.Device
  ordinal::Int
  ROCm(ordinal::Integer=0) = (check(ccall((:ptf_device_init, libptf), Cint, (Int32,), ordinal));
                              new(ordinal))
end

"Device buffer owned by libptf_hip (freed by a finalizer)."
mutable struct PtfArray{T,N}
  ptr::Ptr{Cvoid}
  dims::NTuple{N,Int}
  function PtfArray{T}(dims::NTuple{N,Int}) where {T,N}
    p = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:ptf_buffer_alloc, libptf), Cint, (UInt64, Ref{Ptr{Cvoid}}),
                prod(dims) * sizeof(T), p))
    a = new{T,N}(p[], dims)
    finalizer(x -> ccall((:ptf_buffer_free, libptf), Cint, (Ptr{Cvoid},), x.ptr), a)
    return a
  end
end

mutable struct RFFTPlan
  h::Ptr{Cvoid}
  RFFTPlan(h::Ptr{Cvoid}) = new(h)
  function RFFTPlan(ndim, n::NTuple{3,Int64}, precision)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:ptf_plan_create, libptf), Cint, (Int32, Ref{NTuple{3,Int64}}, Int32,
                Ref{Ptr{Cvoid}}), ndim, n, precision, h))
    p = new(h[])
    finalizer(x -> ccall((:ptf_plan_destroy, libptf), Cint, (Ptr{Cvoid},), x.h), p)
    return p
  end
end

"Layered plan: nlayers fields of the grid transformed over dims 1:ndim (S:678-679, S:865-866)."
function LayeredRFFTPlan(ndim, n::NTuple{3,Int64}, nlayers::Integer, precision)
  h = Ref{Ptr{Cvoid}}(C_NULL)
  check(ccall((:ptf_plan_create_batched, libptf), Cint, (Int32, Ref{NTuple{3,Int64}}, Int64, Int32,
              Ref{Ptr{Cvoid}}), ndim, n, nlayers, precision, h))
  p = RFFTPlan(h[])
  finalizer(x -> ccall((:ptf_plan_destroy, libptf), Cint, (Ptr{Cvoid},), x.h), p)
  return p
end

"dealias!(fh, grid) for a device half spectrum (aliased_fraction 1/3 = the 2/3 rule)."
dealias!(fh::PtfArray, plan::RFFTPlan; aliased_fraction=1/3) =
  check(ccall((:ptf_dealias, libptf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Float64), plan.h, fh.ptr,
              aliased_fraction))

mul!(out::PtfArray, plan::RFFTPlan, a::PtfArray) =
  (check(ccall((:ptf_rfft, libptf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
               plan.h, a.ptr, out.ptr)); out)
ldiv!(out::PtfArray, plan::RFFTPlan, a::PtfArray) =
  (check(ccall((:ptf_irfft, libptf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
               plan.h, a.ptr, out.ptr)); out)

# ---- problem ----------------------------------------------------------------------
mutable struct Clock{T}
  dt::T
  t::T
  step::Int
end

mutable struct DeviceProblem{F}
  h::Ptr{Cvoid}
  ndim::Int
  n::NTuple{3,Int}
  L::NTuple{3,Float64}
  flow::F
  steadyflow::Bool
  clock::Clock{Float64}
  vars::NamedTuple
end

gridpoints_1d(n, L) = [Float64(-big(L) / 2 + j * big(L) / n) for j in 0:n-1]

function points(ndim, n, L)
  x = gridpoints_1d(n[1], L[1])
  ndim == 1 && return (x,)
  y = gridpoints_1d(n[2], L[2])
  ndim == 2 && return ([xi for xi in x, yj in y], [yj for xi in x, yj in y])
  z = gridpoints_1d(n[3], L[3])
  return ([xi for xi in x, yj in y, zk in z], [yj for xi in x, yj in y, zk in z],
          [zk for xi in x, yj in y, zk in z])
end

flowfuncs(f::OneDAdvectingFlow) = (f.u,)
flowfuncs(f::TwoDAdvectingFlow) = (f.u, f.v)
flowfuncs(f::ThreeDAdvectingFlow) = (f.u, f.v, f.w)
ndims_of(f) = length(flowfuncs(f))

function upload_flow!(p::DeviceProblem, t)
  pts = points(p.ndim, p.n, p.L)
  fs = flowfuncs(p.flow)
  fields = [Float64.(p.steadyflow ? f.(pts...) : f.(pts..., t)) .+ zeros(p.n[1:p.ndim]...)
            for f in fs]
  ptrs = [pointer(a) for a in fields]
  while length(ptrs) < 3; push!(ptrs, Ptr{Float64}(C_NULL)); end
  GC.@preserve fields check(ccall((:ptf_set_flow, libptf), Cint,
    (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}), p.h, ptrs[1], ptrs[2], ptrs[3]))
end

"""
    Problem(dev::ROCm, flow; nx=128, Lx=2π, ny=nx, Ly=Lx, nz=nx, Lz=Lx, κ=0.1, η=κ, ι=κ,
            κh=0, nκh=0, dt=0.01, stepper="RK4", T=Float64)

TracerAdvectionDiffusion.Problem (src/traceradvectiondiffusion.jl:143-216) on the
ROCm device.  `T=Float32` selects the mixed-precision form.
"""
function Problem(dev::ROCm, flow; nx=128, Lx=2π, ny=nx, Ly=Lx, nz=nx, Lz=Lx, κ=0.1, η=κ, ι=κ,
                 κh=0.0, nκh=0, dt=0.01, stepper="RK4", T=Float64,
                 order=0, innerK=0.65, outerK=1.0, tol=1e-15,
                 analytic::Union{Nothing,Vector{PtfFlowTerm}}=nothing)
  nd = ndims_of(flow)
  n = (nx, nd >= 2 ? ny : 1, nd >= 3 ? nz : 1)
  L = (Float64(Lx), nd >= 2 ? Float64(Ly) : 1.0, nd >= 3 ? Float64(Lz) : 1.0)
  # filter keywords of FilteredRK4/FilteredETDRK4 ([FF] makefilter); order 0 = its defaults
  d = PtfDesc(nd, T == Float64 ? PTF_FP64 : PTF_MIXED, Int64.(n), L, κ, η, ι, κh, nκh,
              STEPPERS[stepper], dt,
              analytic !== nothing ? PTF_FLOW_ANALYTIC :
              (flow.steadyflow ? PTF_FLOW_STEADY : PTF_FLOW_PER_STEP), order, innerK, outerK, tol)
  h = Ref{Ptr{Cvoid}}(C_NULL)
  check(ccall((:ptf_problem_create, libptf), Cint, (Ref{PtfDesc}, Ref{Ptr{Cvoid}}), d, h))
  vars = (c = zeros(Float64, n[1:nd]...), ch = zeros(ComplexF64, n[1] ÷ 2 + 1, n[2:nd]...))
  p = DeviceProblem(h[], nd, n, L, flow, flow.steadyflow, Clock(Float64(dt), 0.0, 0), vars)
  finalizer(x -> ccall((:ptf_problem_destroy, libptf), Cint, (Ptr{Cvoid},), x.h), p)
  if analytic !== nothing        # `flow` only gives the dimension; the GPU synthesises u, v, w
    p.steadyflow = true
    set_analytic_flow!(p, analytic)
  elseif p.steadyflow
    upload_flow!(p, 0.0)
  end
  return p
end

function set_c!(p::DeviceProblem, c)
  cc = Float64.(c)
  GC.@preserve cc check(ccall((:ptf_set_c, libptf), Cint, (Ptr{Cvoid}, Ptr{Float64}), p.h,
                              pointer(cc)))
  updatevars!(p)
  return nothing
end

function updatevars!(p::DeviceProblem)
  check(ccall((:ptf_updatevars, libptf), Cint, (Ptr{Cvoid}, Ptr{Float64}), p.h, p.vars.c))
  check(ccall((:ptf_get_sol, libptf), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}), p.h, p.vars.ch))
  return nothing
end

function sync_clock!(p::DeviceProblem)
  t = Ref{Float64}(0); s = Ref{Int64}(0)
  check(ccall((:ptf_get_clock, libptf), Cint, (Ptr{Cvoid}, Ref{Float64}, Ref{Int64}), p.h, t, s))
  p.clock.t = t[]; p.clock.step = s[]
end

"""
    stepforward!(prob, diags, nsteps)

FourierFlows' stepforward! with diagnostics; when every diagnostic is `mass` or
`variance` (given as symbols :mass, :variance here), they are recorded on the GPU by
one ptf_step_record call every `freq` steps and returned as (t, steps, values).
"""
function step_record!(p::DeviceProblem, nsteps::Integer, freq::Integer)
  s0 = p.clock.step
  nrec = div(s0 + nsteps, freq) - div(s0, freq)
  t = zeros(max(nrec, 1)); st = zeros(Int64, max(nrec, 1))
  m = zeros(max(nrec, 1)); v = zeros(max(nrec, 1))
  n = Ref{Int64}(0)
  check(ccall((:ptf_step_record, libptf), Cint,
              (Ptr{Cvoid}, Int64, Int64, Int64, Ref{Int64}, Ptr{Float64}, Ptr{Int64}, Ptr{Float64},
               Ptr{Float64}, Ptr{Float64}),
              p.h, nsteps, freq, max(nrec, 1), n, t, st, m, v, C_NULL))
  sync_clock!(p)
  k = n[]
  return (t=t[1:k], steps=st[1:k], mass=m[1:k], variance=v[1:k])
end

"Use a built-in analytic flow (terms synthesised on the GPU at clock.t every step)."
function set_analytic_flow!(p::DeviceProblem, terms::Vector{PtfFlowTerm})
  check(ccall((:ptf_set_analytic_flow, libptf), Cint, (Ptr{Cvoid}, Ptr{PtfFlowTerm}, Int32), p.h,
              terms, length(terms)))
end

function stepforward!(p::DeviceProblem, nsteps::Integer=1)
  if p.steadyflow
    check(ccall((:ptf_step, libptf), Cint, (Ptr{Cvoid}, Int64), p.h, nsteps))
    sync_clock!(p)
  else
    for _ in 1:nsteps
      upload_flow!(p, p.clock.t)
      check(ccall((:ptf_step, libptf), Cint, (Ptr{Cvoid}, Int64), p.h, 1))
      sync_clock!(p)
    end
  end
  return nothing
end

end # module
