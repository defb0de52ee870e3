"""The FourierFlows layer the tracer module plugs into, on the ROCm device.

Mirrors the parts of FourierFlows.jl 0.10.x that PassiveTracerFlows.jl calls
(src/traceradvectiondiffusion.jl, abbreviated S): ``Device``/``GPU()``
(S:505, S:646, S:847), ``OneDGrid/TwoDGrid/ThreeDGrid`` and ``gridpoints``
(S:152, S:177, S:205, S:427, S:437, S:446), ``grid.rfftplan`` with
``mul!``/``ldiv!`` (S:20, S:698-783), ``Problem``/``Clock`` (S:162, S:187,
S:215) and ``stepforward!``.  Julia's ``!`` suffix becomes a trailing ``_``.

Arrays: a physical field is a C-ordered numpy array of shape (nz, ny, nx)
(x fastest: the memory layout of Julia's column-major (nx, ny, nz)); a
spectral field has shape (nz, ny, nx//2+1), complex128.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import call, dptr


class Device:
    pass


class ROCm(Device):
    """One MI355X (HIP device ``ordinal``); every array of the problem lives in its HBM."""

    def __init__(self, ordinal: int = 0):
        self.ordinal = int(ordinal)
        call("ptf_device_init", self.ordinal)

    def __repr__(self):
        return f"ROCm({self.ordinal})"


GPU = ROCm


def _points(n, L):
    start = np.longdouble(-L / 2)
    step = np.longdouble(L / n)
    return (start + np.arange(n, dtype=np.longdouble) * step).astype(np.float64)


def _kscale(n, L):
    dx = L / n
    return (2 * np.pi / dx) / n


class _Grid:
    ndim = 0

    def __init__(self, dev, nx, Lx, ny, Ly, nz, Lz, T, aliased_fraction=1 / 3):
        if not isinstance(dev, ROCm):
            raise TypeError("this grid lives on the ROCm device")
        self.device = dev
        self.T = np.dtype(T).type
        if self.T not in (np.float64, np.float32):
            raise TypeError("T must be Float64 (np.float64) or Float32 (np.float32)")
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.Lx, self.Ly, self.Lz = float(Lx), float(Ly), float(Lz)
        for n in (self.nx, self.ny, self.nz)[: self.ndim]:
            if n % 2:
                raise ValueError("grid sizes must be even")
        self.nkr = self.nx // 2 + 1
        self.nl = self.ny
        self.nm = self.nz
        self.dx = self.Lx / self.nx
        self.dy = self.Ly / self.ny
        self.dz = self.Lz / self.nz
        self.x = _points(self.nx, self.Lx)
        self.y = _points(self.ny, self.Ly) if self.ndim >= 2 else None
        self.z = _points(self.nz, self.Lz) if self.ndim >= 3 else None
        kr = np.arange(self.nkr) * _kscale(self.nx, self.Lx)
        fy = np.fft.fftfreq(self.ny, 1.0 / self.ny) * _kscale(self.ny, self.Ly)
        fz = np.fft.fftfreq(self.nz, 1.0 / self.nz) * _kscale(self.nz, self.Lz)
        if self.ndim == 1:
            self.kr = kr
            self.Krsq = kr ** 2
            self.shape = (self.nx,)
        elif self.ndim == 2:
            self.kr = kr[None, :]
            self.l = fy[:, None]
            self.Krsq = self.kr ** 2 + self.l ** 2
            self.shape = (self.ny, self.nx)
        else:
            self.kr = kr[None, None, :]
            self.l = fy[None, :, None]
            self.m = fz[:, None, None]
            self.Krsq = self.kr ** 2 + self.l ** 2 + self.m ** 2
            self.shape = (self.nz, self.ny, self.nx)
        self.spec_shape = self.shape[:-1] + (self.nkr,)
        self.aliased_fraction = float(aliased_fraction)
        self._plan = None

    @property
    def precision(self):
        return _lib.PTF_FP64 if self.T is np.float64 else _lib.PTF_MIXED

    @property
    def rfftplan(self):
        if self._plan is None:
            self._plan = RFFTPlan(self)
        return self._plan

    def __eltype__(self):
        return self.T


class OneDGrid(_Grid):
    ndim = 1

    def __init__(self, dev, nx=128, Lx=2 * np.pi, T=np.float64, aliased_fraction=1 / 3):
        super().__init__(dev, nx, Lx, 1, 1.0, 1, 1.0, T, aliased_fraction)


class TwoDGrid(_Grid):
    ndim = 2

    def __init__(self, dev, nx=128, Lx=2 * np.pi, ny=None, Ly=None, T=np.float64,
                 aliased_fraction=1 / 3):
        super().__init__(dev, nx, Lx, nx if ny is None else ny, Lx if Ly is None else Ly, 1, 1.0, T,
                         aliased_fraction)


class ThreeDGrid(_Grid):
    ndim = 3

    def __init__(self, dev, nx=128, Lx=2 * np.pi, ny=None, Ly=None, nz=None, Lz=None, T=np.float64,
                 aliased_fraction=1 / 3):
        super().__init__(dev, nx, Lx, nx if ny is None else ny, Lx if Ly is None else Ly,
                         nx if nz is None else nz, Lx if Lz is None else Lz, T, aliased_fraction)


def gridpoints(grid):
    """Coordinate arrays broadcast to the physical shape ([FF] gridpoints)."""
    if grid.ndim == 1:
        return grid.x.copy()
    if grid.ndim == 2:
        Y, X = np.meshgrid(grid.y, grid.x, indexing="ij")
        return X, Y
    Z, Y, X = np.meshgrid(grid.z, grid.y, grid.x, indexing="ij")
    return X, Y, Z


# ---------------------------------------------------------------------------
# device arrays and the rfft plan (mul!/ldiv!)
# ---------------------------------------------------------------------------
class DeviceArray:
    """An HBM buffer owned by libptf_hip, with numpy shape/dtype metadata."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        call("ptf_buffer_alloc", self.nbytes, C.byref(p))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        call("ptf_buffer_upload", self.ptr, a.ctypes.data, self.nbytes)

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        call("ptf_buffer_download", out.ctypes.data, self.ptr, self.nbytes)
        return out

    def __del__(self):
        try:
            if getattr(self, "ptr", None):
                _lib.load().ptf_buffer_free(self.ptr)
                self.ptr = None
        except Exception:
            pass


def variant_flags(occ3=False, par=False, wg=False):
    """PTF_VARIANT_* bits: OCC3 = 1, PAR = 2, WG_LINES = 4 (include/ptf.h)."""
    return (1 if occ3 else 0) | (2 if par else 0) | (4 if wg else 0)


def _devptr(a, shape=None, dtype=None, what="array"):
    """Device address of a DeviceArray after checking it against what the plan expects.

    Only library-owned DeviceArrays are accepted: the kernels index a buffer by the
    plan's sizes, so a too-small, strided or wrongly typed buffer would be read and
    written out of bounds.  (Foreign device tensors are refused for the same reason,
    and because a framework that brings its own HIP runtime into the process makes
    the library refuse to start.)"""
    if not isinstance(a, DeviceArray):
        raise TypeError(f"{what}: expected a DeviceArray (a libptf_hip buffer), got {type(a).__name__}")
    if not a.ptr:
        raise ValueError(f"{what}: the DeviceArray has been freed")
    if shape is not None and tuple(a.shape) != tuple(shape):
        raise ValueError(f"{what}: shape {tuple(a.shape)} does not match the plan's {tuple(shape)}")
    if dtype is not None and a.dtype != np.dtype(dtype):
        raise ValueError(f"{what}: dtype {a.dtype} does not match the plan's {np.dtype(dtype)}")
    return a.ptr


class RFFTPlan:
    """grid.rfftplan: real<->complex FFT over all grid axes on the device.

    ``nlayers > 1`` makes the layered plan of the multi-layer tracer (the
    ``(nx, ny, nlayers)`` transforms over dims 1:2 at S:678-679, S:865-866): arrays
    hold nlayers fields back to back (numpy shape ``(nlayers,) + grid.shape``) and
    every layer is transformed independently in one launch per pass."""

    def __init__(self, grid, nlayers: int = 1):
        self.grid = grid
        self.nlayers = int(nlayers)
        n = (C.c_int64 * 3)(grid.nx, grid.ny, grid.nz)
        h = C.c_void_p()
        if self.nlayers == 1:
            call("ptf_plan_create", grid.ndim, n, grid.precision, C.byref(h))
        else:
            call("ptf_plan_create_batched", grid.ndim, n, self.nlayers, grid.precision, C.byref(h))
        self._h = h.value

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                _lib.load().ptf_plan_destroy(self._h)
                self._h = None
        except Exception:
            pass


def _plan_shapes(plan: RFFTPlan):
    """(physical shape, spectral shape, real dtype, complex dtype) of the plan's arrays."""
    g = plan.grid
    lead = (plan.nlayers,) if plan.nlayers > 1 else ()
    rt = np.dtype(g.T)
    ct = np.dtype(np.complex128 if rt == np.float64 else np.complex64)
    return lead + tuple(g.shape), lead + tuple(g.spec_shape), rt, ct


def mul_(out, plan: RFFTPlan, a):
    """mul!(out, rfftplan, a): unnormalised real-to-complex transform (device arrays).

    Like FFTW's mul!, the arrays must have the plan's sizes (and its precision):
    ValueError otherwise, before anything is launched."""
    rs, ss, rt, ct = _plan_shapes(plan)
    pin = _devptr(a, rs, rt, "input")
    pout = _devptr(out, ss, ct, "output")
    call("ptf_rfft", plan._h, pin, pout)
    return out


def ldiv_(out, plan: RFFTPlan, a):
    """ldiv!(out, rfftplan, a): complex-to-real inverse scaled by 1/(nx ny nz); `a` is preserved.
    Shapes and precision are checked as in mul_."""
    rs, ss, rt, ct = _plan_shapes(plan)
    pin = _devptr(a, ss, ct, "input")
    pout = _devptr(out, rs, rt, "output")
    call("ptf_irfft", plan._h, pin, pout)
    return out


def dealias_(fh, grid, plan: RFFTPlan | None = None):
    """dealias!(fh, grid): zero the aliased wavenumbers of the half spectrum fh (a device
    array; grid.aliased_fraction, default 1/3 = the 2/3 rule).  Pass the layered plan
    for an (nlayers, ...) array."""
    plan = grid.rfftplan if plan is None else plan
    _, ss, _, ct = _plan_shapes(plan)
    call("ptf_dealias", plan._h, _devptr(fh, ss, ct, "fh"), float(grid.aliased_fraction))
    call("ptf_device_synchronize")
    return None


# ---------------------------------------------------------------------------
# Problem / Clock / stepforward!
# ---------------------------------------------------------------------------
class Clock:
    def __init__(self, dt, t=0.0, step=0):
        self.dt = dt
        self.t = t
        self.step = step

    def __repr__(self):
        return f"Clock(dt={self.dt}, t={self.t}, step={self.step})"


class Equation:
    """FourierFlows.Equation(L, calcN!, grid): the diagonal linear operator and the
    nonlinear term; on this device calcN! is executed inside the fused step kernels."""

    def __init__(self, L, calcN, grid, params=None):
        self.L = L
        self.calcN = calcN
        self.grid = grid
        self.params = params
        self.dims = L.shape
        self.T = np.complex128 if grid.T is np.float64 else np.complex64


def tuning_key(prob):
    """Identity of a problem's launch list for the tuning cache."""
    g = prob.grid
    from . import traceradvectiondiffusion as tad
    prec = "fp64" if g.T is np.float64 else "mixed"
    return (f"{g.ndim}D {g.nx}x{g.ny}x{g.nz} {prec} {prob.timestepper} "
            f"flow{tad._flow_kind(prob.params)}")


class Problem:
    """FourierFlows.Problem(eqn, stepper, dt, grid, vars, params) on the ROCm device.

    ``sol`` is device resident; reading ``prob.sol`` downloads a host copy
    (complex128, spectral shape) and assigning it uploads.
    """

    def __init__(self, eqn, stepper, dt, grid, vars, params, **stepperkwargs):
        from . import traceradvectiondiffusion as tad  # the only physics module

        self.eqn = eqn
        self.grid = grid
        self.vars = vars
        self.params = params
        self.stepper = stepper
        st = stepper.upper()
        if st not in _lib.STEPPERS:
            raise ValueError(f"stepper {stepper!r} is not implemented on the ROCm device "
                             "(RK4, ETDRK4, FilteredRK4 and FilteredETDRK4 are)")
        filterkw = None
        if st.startswith("FILTERED"):
            unknown = set(stepperkwargs) - {"order", "innerK", "outerK", "tol"}
            if unknown:
                raise TypeError(f"unknown filter keywords {sorted(unknown)}")
            filterkw = dict(order=4, innerK=0.65, outerK=1.0, tol=1e-15)
            filterkw.update(stepperkwargs)
        elif stepperkwargs:
            raise TypeError(f"stepper {stepper!r} takes no keywords")
        self.clock = Clock(float(dt))
        self.timestepper = st
        self._h = tad._create_handle(self, st, float(dt), filterkw)

    # prob.sol
    @property
    def sol(self):
        out = np.empty(self.grid.spec_shape, np.complex128)
        call("ptf_get_sol", self._h, dptr(out.view(np.float64)))
        return out

    @sol.setter
    def sol(self, value):
        v = np.ascontiguousarray(np.broadcast_to(value, self.grid.spec_shape), dtype=np.complex128)
        call("ptf_set_sol", self._h, dptr(v.view(np.float64)))

    def _sync_clock(self):
        t = C.c_double()
        s = C.c_int64()
        call("ptf_get_clock", self._h, C.byref(t), C.byref(s))
        self.clock.t = t.value
        self.clock.step = s.value

    def stream(self):
        s = C.c_void_p()
        call("ptf_problem_get_stream", self._h, C.byref(s))
        return s.value

    def set_stream(self, hip_stream):
        call("ptf_problem_set_stream", self._h, hip_stream)

    def autotune(self, cache=None):
        """Pick the fastest compiled kernel variant of every launch slot (ptf_autotune).

        With ``cache`` (a JSON file path) the choice is persistent: if the file holds an
        entry for this problem's shape, precision, stepper and flow kind whose slot list
        matches, those variants are applied without timing anything; otherwise the
        autotuner runs and its choice is stored there.  Returns "cache" or "tuned"."""
        if cache is None:
            call("ptf_autotune", self._h)
            return "tuned"
        from passivetracerflows import tuning
        return tuning.apply_or_tune(self, tuning_key(self), cache)

    def set_graph(self, enable: bool):
        """Replay captured 8-step graphs for steady flows (default) or launch directly."""
        call("ptf_problem_set_graph", self._h, 1 if enable else 0)

    def slots(self):
        n = C.c_int32()
        call("ptf_problem_num_slots", self._h, C.byref(n))
        names = []
        for i in range(n.value):
            buf = C.create_string_buffer(256)
            call("ptf_problem_slot_name", self._h, i, buf, 256)
            names.append(buf.value.decode())
        return names

    def time_slot(self, i, reps=20):
        ms = C.c_double()
        call("ptf_problem_time_slot", self._h, i, reps, C.byref(ms))
        return ms.value

    def set_variant(self, i, R, W, occ3=False, par=False, wg=False):
        """Select kernel variant (radix R, width W, 3-waves/SIMD build, input-parallel
        advection, bricked column lines spread over the workgroup) for launch slot i;
        raises PTFError if it was not compiled or does not apply."""
        call("ptf_problem_set_variant", self._h, i, R, W, variant_flags(occ3, par, wg))

    def get_variant(self, i):
        """(R, W, occ3, par, wg) of launch slot i."""
        R, W, f = C.c_int32(), C.c_int32(), C.c_int32()
        call("ptf_problem_get_variant", self._h, i, C.byref(R), C.byref(W), C.byref(f))
        return R.value, W.value, bool(f.value & 1), bool(f.value & 2), bool(f.value & 4)

    def step_profiled(self, nsteps):
        """Advance nsteps steps (as stepforward_) with a HIP event around every launch;
        returns the average duration in ms of each launch slot over those steps."""
        n = C.c_int32()
        call("ptf_problem_num_slots", self._h, C.byref(n))
        out = (C.c_double * n.value)()
        call("ptf_step_profiled", self._h, int(nsteps), out)
        self._sync_clock()
        return list(out)

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                _lib.load().ptf_problem_destroy(self._h)
                self._h = None
        except Exception:
            pass


def stepforward_(prob: Problem, nsteps=1, nsteps2=None):
    """stepforward!(prob[, nsteps]): advance, then clock.t += dt and clock.step += 1 per step.
    stepforward!(prob, diags, nsteps): after each step, increment!(diags).

    A time-varying user flow is evaluated on the host at clock.t before each
    step (it is frozen within the step, as calcN! at S:701, S:718, S:737).
    """
    from . import traceradvectiondiffusion as tad

    if nsteps2 is not None:
        from .output import stepforward_diags_
        return stepforward_diags_(prob, nsteps, nsteps2)
    nsteps = int(nsteps)
    if nsteps < 0:
        raise ValueError("nsteps must be >= 0")
    if tad._is_host_time_varying(prob.params):
        for _ in range(nsteps):
            tad._upload_flow_at(prob, prob.clock.t)
            call("ptf_step", prob._h, 1)
            prob._sync_clock()
    else:
        call("ptf_step", prob._h, nsteps)
        prob._sync_clock()
    return None


def step_until_(prob: Problem, tfinal: float):
    """step_until!(prob, tfinal): step while clock.t < tfinal (tfinal a multiple of dt)."""
    n = int(round((tfinal - prob.clock.t) / prob.clock.dt))
    if n < 0:
        raise ValueError("tfinal is before the current time")
    stepforward_(prob, n)
