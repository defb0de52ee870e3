"""TracerAdvectionDiffusion on the ROCm device: the drop-in for the reference module.

Mirrors PassiveTracerFlows.jl v0.9.1 ``src/traceradvectiondiffusion.jl`` (S):

* ``OneDAdvectingFlow/TwoDAdvectingFlow/ThreeDAdvectingFlow``, ``noflow``  -- S:29-106
* ``Problem(dev, flow; nx, Lx, ny, Ly, nz, Lz, κ, η, ι, dt, stepper, T)``   -- S:143-216
* ``ConstDiffSteadyFlowParams``, ``ConstDiffTimeVaryingFlowParams``          -- S:256-452
* ``Equation(params, grid)``                                              -- S:502-554
* ``Vars(grid)``                                                          -- S:572-668
* ``calcN_`` (calcN!)                                                     -- S:695-786
* ``updatevars_`` (updatevars!), ``set_c_`` (set_c!)                      -- S:815-872

Keyword names follow the reference (``κ``, ``η``, ``ι``, ``κh``, ``nκh``);
ASCII spellings ``kappa``, ``eta``, ``iota``, ``kappah``, ``nkappah`` are
accepted too.  ``T=np.float32`` selects the mixed-precision form (FP32 bulk,
FP64 setup coefficients and diagnostics).  The stepping, transforms and
calcN! all run in libptf_hip on the GPU; there is no CPU path.

Additions without a reference counterpart: ``SpectralStreamfunctionFlow``
and the analytic flows of ``passivetracerflows.flows`` (built-in flows
synthesised on the GPU), ``mass``/``variance`` (FP64 diagnostics).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import call, dptr
from .fourierflows import (Equation as _FFEquation, OneDGrid, Problem as _FFProblem, ROCm,
                           ThreeDGrid, TwoDGrid, gridpoints, stepforward_)

__all__ = [
    "Problem", "set_c_", "updatevars_", "OneDAdvectingFlow", "TwoDAdvectingFlow",
    "ThreeDAdvectingFlow", "SpectralStreamfunctionFlow", "noflow", "Vars", "Equation",
    "ConstDiffSteadyFlowParams", "ConstDiffTimeVaryingFlowParams", "calcN_", "mass", "variance",
]


def noflow(*args):
    """Default flow component: 0.0 everywhere (S:31)."""
    return 0.0


# ---------------------------------------------------------------------------
# advecting flows (S:29-106)
# ---------------------------------------------------------------------------
class AbstractAdvectingFlow:
    pass


class OneDAdvectingFlow(AbstractAdvectingFlow):
    def __init__(self, u=noflow, steadyflow=True):
        self.u = u
        self.steadyflow = bool(steadyflow)


class TwoDAdvectingFlow(AbstractAdvectingFlow):
    def __init__(self, u=noflow, v=noflow, steadyflow=True):
        self.u, self.v = u, v
        self.steadyflow = bool(steadyflow)


class ThreeDAdvectingFlow(AbstractAdvectingFlow):
    def __init__(self, u=noflow, v=noflow, w=noflow, steadyflow=True):
        self.u, self.v, self.w = u, v, w
        self.steadyflow = bool(steadyflow)


class SpectralStreamfunctionFlow(AbstractAdvectingFlow):
    """2D flow u = -d_y psi, v = d_x psi with psi_hat(k, t) = psi0_hat(k) exp(-i omega(k) t).

    ``psih0`` is the rfft half spectrum (ny, nx//2+1) and ``omega`` a real
    array of the same shape; the flow is synthesised on the GPU at clock.t
    once per step.
    """

    steadyflow = False

    def __init__(self, psih0, omega):
        self.psih0 = np.ascontiguousarray(psih0, dtype=np.complex128)
        self.omega = np.ascontiguousarray(omega, dtype=np.float64)


# ---------------------------------------------------------------------------
# params (S:256-452)
# ---------------------------------------------------------------------------
class _Params:
    steady = True
    ndim = 0

    def __init__(self, kappa, eta, iota, kappah, nkappah, flows):
        self.κ = float(kappa)
        self.η = float(eta)
        self.ι = float(iota)
        self.κh = float(kappah)
        self.nκh = int(nkappah)
        names = ("u", "v", "w")
        for n, f in zip(names, flows):
            setattr(self, n, f)

    kappa = property(lambda s: s.κ)
    eta = property(lambda s: s.η)
    iota = property(lambda s: s.ι)
    kappah = property(lambda s: s.κh)
    nkappah = property(lambda s: s.nκh)


class ConstDiffSteadyFlowParams1D(_Params):
    ndim = 1


class ConstDiffSteadyFlowParams2D(_Params):
    ndim = 2


class ConstDiffSteadyFlowParams3D(_Params):
    ndim = 3


class ConstDiffTimeVaryingFlowParams1D(_Params):
    ndim = 1
    steady = False


class ConstDiffTimeVaryingFlowParams2D(_Params):
    ndim = 2
    steady = False


class ConstDiffTimeVaryingFlowParams3D(_Params):
    ndim = 3
    steady = False


class ConstDiffSpectralFlowParams2D(_Params):
    ndim = 2
    steady = False


class ConstDiffAnalyticFlowParams(_Params):
    """Diffusivities plus an AnalyticAdvectingFlow (synthesised on the GPU)."""
    steady = False

    def __init__(self, kappa, eta, iota, kappah, nkappah, flow):
        super().__init__(kappa, eta, iota, kappah, nkappah, ())
        self.flow = flow
        self.ndim = flow.ndim


def _sample(f, pts, shape):
    """u.(x, y, ...) on the grid: vectorised call, broadcast to the grid shape."""
    try:
        val = f(*pts)
    except (TypeError, ValueError):
        val = np.vectorize(f, otypes=[np.float64])(*pts)
    return np.ascontiguousarray(np.broadcast_to(np.asarray(val, dtype=np.float64), shape))


def ConstDiffSteadyFlowParams(*args):
    """Steady-flow params; the flow functions are sampled once on the grid (S:426-452).

    Forms (as the reference): (κ, u, grid1D), (κ, κh, nκh, u, grid1D),
    (κ, η, u, v, grid2D), (κ, η, κh, nκh, u, v, grid2D),
    (κ, η, ι, u, v, w, grid3D), (κ, η, ι, κh, nκh, u, v, w, grid3D).
    """
    grid = args[-1]
    a = args[:-1]
    if isinstance(grid, OneDGrid):
        if len(a) == 2:
            kappa, u = a
            kappah, nkappah = 0.0 * kappa, 0
        else:
            kappa, kappah, nkappah, u = a
        x = gridpoints(grid)
        return ConstDiffSteadyFlowParams1D(kappa, kappa, kappa, kappah, nkappah,
                                           (_sample(u, (x,), grid.shape),))
    if isinstance(grid, TwoDGrid):
        if len(a) == 4:
            kappa, eta, u, v = a
            kappah, nkappah = 0.0 * kappa, 0
        else:
            kappa, eta, kappah, nkappah, u, v = a
        X, Y = gridpoints(grid)
        return ConstDiffSteadyFlowParams2D(kappa, eta, kappa, kappah, nkappah,
                                           (_sample(u, (X, Y), grid.shape),
                                            _sample(v, (X, Y), grid.shape)))
    if isinstance(grid, ThreeDGrid):
        if len(a) == 6:
            kappa, eta, iota, u, v, w = a
            kappah, nkappah = 0.0 * kappa, 0
        else:
            kappa, eta, iota, kappah, nkappah, u, v, w = a
        X, Y, Z = gridpoints(grid)
        return ConstDiffSteadyFlowParams3D(kappa, eta, iota, kappah, nkappah,
                                           tuple(_sample(f, (X, Y, Z), grid.shape)
                                                 for f in (u, v, w)))
    raise TypeError("last argument must be a grid")


def ConstDiffTimeVaryingFlowParams(*args):
    """Time-varying params (S:334-348): (κ, u), (κ, η, u, v), (κ, η, ι, u, v, w);
    κh = 0, nκh = 0.  The functions take (x[, y[, z]], t)."""
    if len(args) == 2:
        kappa, u = args
        return ConstDiffTimeVaryingFlowParams1D(kappa, kappa, kappa, 0.0 * kappa, 0, (u,))
    if len(args) == 4:
        kappa, eta, u, v = args
        return ConstDiffTimeVaryingFlowParams2D(kappa, eta, kappa, 0.0 * kappa, 0, (u, v))
    if len(args) == 6:
        kappa, eta, iota, u, v, w = args
        return ConstDiffTimeVaryingFlowParams3D(kappa, eta, iota, 0.0 * kappa, 0, (u, v, w))
    raise TypeError("ConstDiffTimeVaryingFlowParams takes 2, 4 or 6 arguments")


def ConstDiffTimeVaryingFlowParamsFull(ndim, kappa, eta, iota, kappah, nkappah, *flows):
    """The struct constructors ConstDiffTimeVaryingFlowParams{1,2,3}D(κ, [η, [ι,]] κh, nκh, u...)."""
    cls = {1: ConstDiffTimeVaryingFlowParams1D, 2: ConstDiffTimeVaryingFlowParams2D,
           3: ConstDiffTimeVaryingFlowParams3D}[ndim]
    return cls(kappa, eta, iota, kappah, nkappah, flows)


def ConstDiffSpectralFlowParams(kappa, eta, kappah, nkappah, flow: SpectralStreamfunctionFlow):
    p = ConstDiffSpectralFlowParams2D(kappa, eta, kappa, kappah, nkappah, ())
    p.flow = flow
    return p


# ---------------------------------------------------------------------------
# Vars (S:572-668)
# ---------------------------------------------------------------------------
class Vars:
    """Physical c (host copy, filled by updatevars_) and its transform ch."""

    def __init__(self, grid):
        self.grid = grid
        self.c = np.zeros(grid.shape)
        self.ch = np.zeros(grid.spec_shape, np.complex128)


Vars1D = Vars2D = Vars3D = Vars


# ---------------------------------------------------------------------------
# Equation (S:502-554)
# ---------------------------------------------------------------------------
def _linear_operator(params, grid):
    if grid.ndim == 1:
        L = -params.κ * grid.kr ** 2 - params.κh * (grid.kr ** 2) ** params.nκh
    elif grid.ndim == 2:
        L = -params.κ * grid.kr ** 2 - params.η * grid.l ** 2 - params.κh * grid.Krsq ** params.nκh
    else:
        L = (-params.κ * grid.kr ** 2 - params.η * grid.l ** 2 - params.ι * grid.m ** 2
             - params.κh * grid.Krsq ** params.nκh)
    return np.ascontiguousarray(np.broadcast_to(L, grid.spec_shape), dtype=np.float64)


def Equation(params, grid):
    """FourierFlows.Equation(L, calcN!, grid) with L = -κkr² - ηl² - ιm² - κh Krsq^nκh."""
    return _FFEquation(_linear_operator(params, grid), calcN_, grid, params)


# ---------------------------------------------------------------------------
# Problem (S:143-216)
# ---------------------------------------------------------------------------
def _kw(kw, greek, ascii_, default):
    if greek in kw:
        return kw.pop(greek)
    if ascii_ in kw:
        return kw.pop(ascii_)
    return default


def Problem(dev, advecting_flow, **kw):
    """TracerAdvectionDiffusion.Problem(dev, advecting_flow; nx=128, Lx=2π, ny=nx, Ly=Lx,
    nz=nx, Lz=Lx, κ=0.1, η=κ, ι=κ, dt=0.01, stepper="RK4", T=Float64)."""
    nx = kw.pop("nx", 128)
    Lx = kw.pop("Lx", 2 * np.pi)
    dt = kw.pop("dt", 0.01)
    stepper = kw.pop("stepper", "RK4")
    T = kw.pop("T", np.float64)
    kappa = _kw(kw, "κ", "kappa", 0.1)
    from .flows import AnalyticAdvectingFlow
    if isinstance(advecting_flow, AnalyticAdvectingFlow):
        nd = advecting_flow.ndim
        ny = kw.pop("ny", nx) if nd >= 2 else 1
        Ly = kw.pop("Ly", Lx) if nd >= 2 else 1.0
        nz = kw.pop("nz", nx) if nd == 3 else 1
        Lz = kw.pop("Lz", Lx) if nd == 3 else 1.0
        eta = _kw(kw, "η", "eta", kappa) if nd >= 2 else kappa
        iota = _kw(kw, "ι", "iota", kappa) if nd == 3 else kappa
        kappah = _kw(kw, "κh", "kappah", 0.0)
        nkappah = _kw(kw, "nκh", "nkappah", 0)
        if kw:
            raise TypeError(f"unexpected keywords: {sorted(kw)}")
        grid = (OneDGrid(dev, nx=nx, Lx=Lx, T=T) if nd == 1 else
                TwoDGrid(dev, nx=nx, Lx=Lx, ny=ny, Ly=Ly, T=T) if nd == 2 else
                ThreeDGrid(dev, nx=nx, Lx=Lx, ny=ny, Ly=Ly, nz=nz, Lz=Lz, T=T))
        params = ConstDiffAnalyticFlowParams(kappa, eta, iota, kappah, nkappah, advecting_flow)
    elif isinstance(advecting_flow, OneDAdvectingFlow):
        if kw:
            raise TypeError(f"unexpected keywords for a 1D problem: {sorted(kw)}")
        grid = OneDGrid(dev, nx=nx, Lx=Lx, T=T)
        params = (ConstDiffSteadyFlowParams(kappa, advecting_flow.u, grid)
                  if advecting_flow.steadyflow else
                  ConstDiffTimeVaryingFlowParams(kappa, advecting_flow.u))
    elif isinstance(advecting_flow, (TwoDAdvectingFlow, SpectralStreamfunctionFlow)):
        ny = kw.pop("ny", nx)
        Ly = kw.pop("Ly", Lx)
        eta = _kw(kw, "η", "eta", kappa)
        kappah = _kw(kw, "κh", "kappah", 0.0)
        nkappah = _kw(kw, "nκh", "nkappah", 0)
        if kw:
            raise TypeError(f"unexpected keywords: {sorted(kw)}")
        grid = TwoDGrid(dev, nx=nx, Lx=Lx, ny=ny, Ly=Ly, T=T)
        if isinstance(advecting_flow, SpectralStreamfunctionFlow):
            params = ConstDiffSpectralFlowParams(kappa, eta, kappah, nkappah, advecting_flow)
        elif advecting_flow.steadyflow:
            params = ConstDiffSteadyFlowParams(kappa, eta, kappah, nkappah,
                                               advecting_flow.u, advecting_flow.v, grid)
        else:
            params = ConstDiffTimeVaryingFlowParamsFull(2, kappa, eta, kappa, kappah, nkappah,
                                                        advecting_flow.u, advecting_flow.v)
    elif isinstance(advecting_flow, ThreeDAdvectingFlow):
        ny = kw.pop("ny", nx)
        Ly = kw.pop("Ly", Lx)
        nz = kw.pop("nz", nx)
        Lz = kw.pop("Lz", Lx)
        eta = _kw(kw, "η", "eta", kappa)
        iota = _kw(kw, "ι", "iota", kappa)
        if kw:
            raise TypeError(f"unexpected keywords: {sorted(kw)}")
        grid = ThreeDGrid(dev, nx=nx, Lx=Lx, ny=ny, Ly=Ly, nz=nz, Lz=Lz, T=T)
        params = (ConstDiffSteadyFlowParams(kappa, eta, iota, advecting_flow.u, advecting_flow.v,
                                            advecting_flow.w, grid)
                  if advecting_flow.steadyflow else
                  ConstDiffTimeVaryingFlowParams(kappa, eta, iota, advecting_flow.u,
                                                 advecting_flow.v, advecting_flow.w))
    else:
        raise TypeError("advecting_flow must be a One/Two/ThreeDAdvectingFlow")
    vars = Vars(grid)
    equation = Equation(params, grid)
    return _FFProblem(equation, stepper, dt, grid, vars, params)


# ---------------------------------------------------------------------------
# device problem creation and flow upload (called by fourierflows.Problem)
# ---------------------------------------------------------------------------
def _is_host_time_varying(params):
    return (not params.steady) and not isinstance(params, (ConstDiffSpectralFlowParams2D,
                                                           ConstDiffAnalyticFlowParams))


def _flow_kind(params):
    if isinstance(params, ConstDiffSpectralFlowParams2D):
        return _lib.PTF_FLOW_SPECTRAL
    if isinstance(params, ConstDiffAnalyticFlowParams):
        return _lib.PTF_FLOW_ANALYTIC
    return _lib.PTF_FLOW_STEADY if params.steady else _lib.PTF_FLOW_PER_STEP


def _desc(params, grid, stepper, dt, filterkw=None):
    d = _lib.ptf_desc()
    d.ndim = grid.ndim
    d.precision = grid.precision
    for i, (n, L) in enumerate(((grid.nx, grid.Lx), (grid.ny, grid.Ly), (grid.nz, grid.Lz))):
        d.n[i] = n if i < grid.ndim else 1
        d.L[i] = L if i < grid.ndim else 1.0
    d.kappa = params.κ
    d.eta = params.η
    d.iota = params.ι
    d.kappah = params.κh
    d.nkappah = params.nκh
    d.stepper = _lib.STEPPERS[stepper.upper()]
    d.dt = dt
    d.flow_kind = _flow_kind(params)
    if filterkw:
        # [FF] makefilter keywords (defaults order=4, innerK=0.65, outerK=1, tol=1e-15)
        d.filter_order = int(filterkw.get("order", 4))
        d.filter_innerK = float(filterkw.get("innerK", 0.65))
        d.filter_outerK = float(filterkw.get("outerK", 1.0))
        d.filter_tol = float(filterkw.get("tol", 1e-15))
    return d


def _upload_fields(h, fields):
    arrs = [np.ascontiguousarray(f, dtype=np.float64) for f in fields]
    ptrs = [dptr(a) for a in arrs] + [None] * (3 - len(arrs))
    call("ptf_set_flow", h, *ptrs)


def _create_handle(prob, stepper, dt, filterkw=None):
    params, grid = prob.params, prob.grid
    d = _desc(params, grid, stepper, dt, filterkw)
    h = C.c_void_p()
    call("ptf_problem_create", C.byref(d), C.byref(h))
    h = h.value
    try:
        if isinstance(params, ConstDiffSpectralFlowParams2D):
            f = params.flow
            if f.psih0.shape != grid.spec_shape or f.omega.shape != grid.spec_shape:
                raise ValueError("psih0/omega must have the spectral shape of the grid")
            call("ptf_set_spectral_flow", h, dptr(f.psih0.view(np.float64)), dptr(f.omega))
        elif isinstance(params, ConstDiffAnalyticFlowParams):
            if params.flow.ndim != grid.ndim:
                raise ValueError("the analytic flow's dimension differs from the grid's")
            terms = (_lib.ptf_flow_term * len(params.flow.terms))(
                *[t.ctype() for t in params.flow.terms])
            call("ptf_set_analytic_flow", h, terms, len(params.flow.terms))
        elif params.steady:
            _upload_fields(h, [getattr(params, n) for n in ("u", "v", "w")[: grid.ndim]])
        prob.vars._prob = prob
    except Exception:
        _lib.load().ptf_problem_destroy(h)
        raise
    return h


def _upload_flow_at(prob, t):
    """Evaluate the user's u(x, ..., t) at clock.t and upload (S:701, S:718, S:737)."""
    grid, params = prob.grid, prob.params
    pts = gridpoints(grid)
    pts = (pts,) if grid.ndim == 1 else tuple(pts)
    fns = [params.u, getattr(params, "v", None), getattr(params, "w", None)][: grid.ndim]
    fields = [_sample(lambda *a, f=f: f(*a, t), pts, grid.shape) for f in fns]
    _upload_fields(prob._h, fields)


# ---------------------------------------------------------------------------
# set_c!, updatevars!, calcN! (S:695-872)
# ---------------------------------------------------------------------------
def set_c_(prob, c):
    """set_c!(prob, c): sol = R2C(c) on the device, then updatevars!."""
    c = np.ascontiguousarray(np.broadcast_to(np.asarray(c, dtype=np.float64), prob.grid.shape))
    call("ptf_set_c", prob._h, dptr(c))
    updatevars_(prob)


def updatevars_(prob):
    """updatevars!(prob): vars.ch = sol; vars.c = C2R(sol) (copied to the host)."""
    out = np.empty(prob.grid.shape)
    call("ptf_updatevars", prob._h, dptr(out))
    prob.vars.c = out
    prob.vars.ch = prob.sol
    return None


def calcN_(N, sol, t, clock, vars, params, grid):
    """calcN!(N, sol, t, clock, vars, params, grid): N = -R2C(u·∇C2R(sol)) with the flow
    at clock.t (host arrays in and out; evaluated by the device kernels)."""
    prob = vars._prob
    if _is_host_time_varying(params):
        _upload_flow_at(prob, clock.t)
    s = np.ascontiguousarray(np.broadcast_to(sol, grid.spec_shape), dtype=np.complex128)
    out = np.empty(grid.spec_shape, np.complex128)
    call("ptf_calcN", prob._h, dptr(s.view(np.float64)), dptr(out.view(np.float64)))
    N[...] = out
    return None


def mass(prob):
    """Σ c over the grid (= Re ĉ at k = 0), FP64."""
    m = C.c_double()
    call("ptf_diagnostics", prob._h, C.byref(m), None)
    return m.value


def variance(prob):
    """mean((c - mean c)²) by a compensated FP64 Parseval sum."""
    m = C.c_double()
    v = C.c_double()
    call("ptf_diagnostics", prob._h, C.byref(m), C.byref(v))
    return v.value


# Julia-style aliases
set_c = set_c_
updatevars = updatevars_
stepforward = stepforward_
