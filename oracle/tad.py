"""FP64 CPU restatement of the TracerAdvectionDiffusion time step.

TEST INFRASTRUCTURE ONLY.  This module is the oracle that the HIP library is
checked against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product path
(``passivetracerflows`` + ``libptf_hip``) never calls it.

Array conventions
-----------------
Julia arrays are column-major with x fastest.  Here a physical field is a
C-ordered numpy array of shape ``(nz, ny, nx)`` (2D: ``(ny, nx)``, 1D:
``(nx,)``), which is the same memory layout.  A spectral field is the
``rfftn`` half-spectrum over all axes, shape ``(nz, ny, nx//2+1)``.

What each function follows (reference = PassiveTracerFlows.jl v0.9.1,
``src/traceradvectiondiffusion.jl`` abbreviated ``S``; the grid, FFT-plan and
timestepper arithmetic lives in FourierFlows.jl 0.10.x, which is not vendored
in the reference; its published algorithm is restated where marked [FF]):

* ``Grid``            -- [FF] ``OneDGrid/TwoDGrid/ThreeDGrid`` as called at S:152,177,205:
                         x_j = -L/2 + j L/n (range arithmetic, rounded once),
                         kr = rfftfreq (Nyquist +n/2), l, m = fftfreq (Nyquist -n/2).
* ``rfft``/``irfft``  -- the plan interface ``mul!``/``ldiv!`` at S:698-783, S:818, S:847:
                         forward unnormalised, inverse scaled by 1/(nx ny nz); the inverse
                         keeps only the real part of the x-DC and x-Nyquist columns (the
                         c2r convention of FFTW and pocketfft alike).
* ``linear_operator`` -- ``Equation`` S:502-554.
* ``calcN``           -- ``calcN!`` S:695-786 (advective form, no dealiasing, flow at clock.t).
* ``Problem.step``    -- [FF] ``RK4TimeStepper``/``ETDRK4TimeStepper`` ``stepforward!``:
                         stage algebra, RHS sum order and clock update after the step.
* ``etd_coefficients``-- [FF] ``getetdcoeffs(dt, L; ncirc=32, rcirc=1)`` and ``getexpLs``.
* ``makefilter``      -- [FF] ``makefilter(grid; order=4, innerK=0.65, outerK=1, tol=1e-15)``
                         for ``FilteredRK4``/``FilteredETDRK4`` (the MQG-coupled problem's
                         stepper, S:228): sol = filter * (update) after every step.
* ``aliased_ranges``/``dealias`` -- [FF] ``getaliasedwavenumbers`` and ``dealias!``
                         (aliased_fraction = 1/3: the 2/3 rule).
* ``set_c``/``updatevars`` -- S:844-852, S:815-821.

Pinning: tests/test_oracle_pins.py checks every function above against closed
forms (brute-force DFT, RK4 amplification polynomial, exp(L dt), mpmath
phi-functions, the reference's own analytic-solution tests at their
tolerances).
"""
from __future__ import annotations

import math
import os

import numpy as np
import scipy.fft as sfft

_WORKERS = int(os.environ.get("PTF_ORACLE_WORKERS", str(min(16, os.cpu_count() or 1))))


def set_workers(n: int) -> None:
    """Number of pocketfft threads used by the oracle transforms."""
    global _WORKERS
    _WORKERS = max(1, int(n))


# ----------------------------------------------------------------------------
# grid  ([FF] domains.jl, called at S:152, S:177, S:205)
# ----------------------------------------------------------------------------
def gridpoints_1d(n: int, L: float) -> np.ndarray:
    """x_j = -L/2 + j*(L/n), j = 0..n-1, evaluated exactly and rounded once.

    Julia's ``range(-L/2, step=L/n, length=n)`` keeps start and step in
    double-double precision, so each point is the correctly rounded value of
    start + j*step.  With n a power of two and start/step doubles the sum is
    exact in 80-bit long double, so one final rounding reproduces it.
    """
    start = np.longdouble(-L / 2)
    step = np.longdouble(L / n)
    j = np.arange(n, dtype=np.longdouble)
    return (start + j * step).astype(np.float64)


def rfftfreq_k(n: int, L: float) -> np.ndarray:
    """kr = rfftfreq(n, 2pi/dx) = [0..n/2] * 2pi/L."""
    dx = L / n
    fs = 2 * np.pi / dx
    return np.arange(n // 2 + 1, dtype=np.float64) * (fs / n)


def fftfreq_k(n: int, L: float) -> np.ndarray:
    """l = fftfreq(n, 2pi/dy) = [0..n/2-1, -n/2..-1] * 2pi/L (Nyquist negative)."""
    dx = L / n
    fs = 2 * np.pi / dx
    idx = np.concatenate([np.arange(0, n // 2), np.arange(-(n // 2), 0)]).astype(np.float64)
    return idx * (fs / n)


class Grid:
    """1D/2D/3D periodic grid; numpy axis order is (z, y, x)."""

    def __init__(self, ndim, nx, Lx=2 * np.pi, ny=None, Ly=None, nz=None, Lz=None):
        self.ndim = int(ndim)
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        Ly = Lx if Ly is None else Ly
        Lz = Lx if Lz is None else Lz
        for n in (nx, ny, nz)[: self.ndim]:
            if n % 2:
                raise ValueError("grid sizes must be even")
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.Lx, self.Ly, self.Lz = float(Lx), float(Ly), float(Lz)
        self.nkr = self.nx // 2 + 1
        self.x = gridpoints_1d(self.nx, self.Lx)
        self.y = gridpoints_1d(self.ny, self.Ly) if ndim >= 2 else None
        self.z = gridpoints_1d(self.nz, self.Lz) if ndim >= 3 else None
        kr = rfftfreq_k(self.nx, self.Lx)
        if self.ndim == 1:
            self.shape = (self.nx,)
            self.kr = kr
            self.l = np.zeros(1)
            self.m = np.zeros(1)
            self.Krsq = kr ** 2
        elif self.ndim == 2:
            self.shape = (self.ny, self.nx)
            self.kr = kr[None, :]
            self.l = fftfreq_k(self.ny, self.Ly)[:, None]
            self.m = np.zeros((1, 1))
            self.Krsq = self.kr ** 2 + self.l ** 2
        else:
            self.shape = (self.nz, self.ny, self.nx)
            self.kr = kr[None, None, :]
            self.l = fftfreq_k(self.ny, self.Ly)[None, :, None]
            self.m = fftfreq_k(self.nz, self.Lz)[:, None, None]
            self.Krsq = self.kr ** 2 + self.l ** 2 + self.m ** 2
        self.spec_shape = self.shape[:-1] + (self.nkr,)
        self.npoints = int(np.prod(self.shape))

    def gridpoints(self):
        """Coordinate arrays broadcast to the physical shape ([FF] gridpoints)."""
        if self.ndim == 1:
            return (self.x.copy(),)
        if self.ndim == 2:
            Y, X = np.meshgrid(self.y, self.x, indexing="ij")
            return X, Y
        Z, Y, X = np.meshgrid(self.z, self.y, self.x, indexing="ij")
        return X, Y, Z


# ----------------------------------------------------------------------------
# plans: mul!/ldiv! (S:698-783, S:818, S:847)
# ----------------------------------------------------------------------------
def rfft(grid: Grid, c: np.ndarray) -> np.ndarray:
    """Unnormalised real-to-complex DFT over all axes (``mul!(out, rfftplan, c)``)."""
    return sfft.rfftn(np.asarray(c, dtype=np.float64), axes=tuple(range(grid.ndim)), workers=_WORKERS)


def irfft(grid: Grid, ch: np.ndarray) -> np.ndarray:
    """Normalised complex-to-real inverse DFT (``ldiv!(out, rfftplan, ch)``).

    Complex inverse along the non-x axes, then a 1D c2r along x that uses only
    the real parts of the kx=0 and kx=nx/2 entries.
    """
    return sfft.irfftn(np.asarray(ch, dtype=np.complex128), s=grid.shape,
                       axes=tuple(range(grid.ndim)), workers=_WORKERS)


# ----------------------------------------------------------------------------
# linear operator  (Equation, S:502-554)
# ----------------------------------------------------------------------------
def linear_operator(grid: Grid, kappa, eta=None, iota=None, kappah=0.0, nkappah=0):
    """L = -kappa kr^2 - eta l^2 - iota m^2 - kappah Krsq^nkappah (1D: -kappa kr^2 - kappah (kr^2)^n).

    ``x**0 == 1`` also for x == 0 (as Julia's ``0.0^0``), so nkappah = 0 with
    kappah != 0 damps every mode by kappah.
    """
    eta = kappa if eta is None else eta
    iota = kappa if iota is None else iota
    if grid.ndim == 1:
        L = -kappa * grid.kr ** 2 - kappah * (grid.kr ** 2) ** nkappah
    elif grid.ndim == 2:
        L = -kappa * grid.kr ** 2 - eta * grid.l ** 2 - kappah * grid.Krsq ** nkappah
    else:
        L = (-kappa * grid.kr ** 2 - eta * grid.l ** 2 - iota * grid.m ** 2
             - kappah * grid.Krsq ** nkappah)
    return np.broadcast_to(L, grid.spec_shape).astype(np.float64).copy()


# ----------------------------------------------------------------------------
# calcN!  (S:695-786)
# ----------------------------------------------------------------------------
def calcN(grid: Grid, sol: np.ndarray, flow) -> np.ndarray:
    """N = R2C(-u C2R(i kr sol) - v C2R(i l sol) - w C2R(i m sol)); flow = (u[, v[, w]]).

    Evaluation order ((-u*cx) - v*cy) - w*cz as S:781.  ``flow`` holds physical
    arrays (steady fields, or the time-varying flow already evaluated at clock.t).
    """
    cx = irfft(grid, 1j * grid.kr * sol)
    acc = -flow[0] * cx
    if grid.ndim >= 2:
        cy = irfft(grid, 1j * grid.l * sol)
        acc = acc - flow[1] * cy
    if grid.ndim >= 3:
        cz = irfft(grid, 1j * grid.m * sol)
        acc = acc - flow[2] * cz
    return rfft(grid, acc)


# ----------------------------------------------------------------------------
# ETDRK4 coefficients  ([FF] getetdcoeffs / getexpLs)
# ----------------------------------------------------------------------------
def etd_coefficients(dt: float, L: np.ndarray, ncirc: int = 32, rcirc: float = 1.0):
    """Kassam-Trefethen contour means: returns (expLdt, expLdt2, zeta, alpha, beta, Gamma).

    z = dt L + rcirc exp(2 pi i (j + 1/2)/ncirc); each coefficient is dt times
    the real part of the mean over the ncirc contour points of
      zeta : (e^{z/2} - 1)/z
      alpha: (-4 - z + e^z (4 - 3z + z^2))/z^3
      beta :  (2 + z + e^z (-2 + z))/z^3
      Gamma: (-4 - 3z - z^2 + e^z (4 - z))/z^3
    """
    j = np.arange(ncirc, dtype=np.float64) + 0.5
    circ = rcirc * np.exp(2j * np.pi / ncirc * j)
    zeta = np.zeros(L.shape)
    alpha = np.zeros(L.shape)
    beta = np.zeros(L.shape)
    gamma = np.zeros(L.shape)
    flat = (dt * L).reshape(-1)
    # process in chunks to bound memory
    out = [a.reshape(-1) for a in (zeta, alpha, beta, gamma)]
    chunk = 1 << 16
    for s in range(0, flat.size, chunk):
        zc = flat[s:s + chunk, None] + circ[None, :]
        ez = np.exp(zc)
        ez2 = np.exp(zc / 2)
        z2 = zc * zc
        z3 = z2 * zc
        out[0][s:s + chunk] = dt * np.real(np.mean((ez2 - 1) / zc, axis=1))
        out[1][s:s + chunk] = dt * np.real(np.mean((-4 - zc + ez * (4 - 3 * zc + z2)) / z3, axis=1))
        out[2][s:s + chunk] = dt * np.real(np.mean((2 + zc + ez * (-2 + zc)) / z3, axis=1))
        out[3][s:s + chunk] = dt * np.real(np.mean((-4 - 3 * zc - z2 + ez * (4 - zc)) / z3, axis=1))
    expLdt = np.exp(dt * L)
    expLdt2 = np.exp(dt * L / 2)
    return expLdt, expLdt2, zeta, alpha, beta, gamma


# ----------------------------------------------------------------------------
# filter and dealiasing  ([FF] utils.jl makefilter; domains.jl getaliasedwavenumbers)
# ----------------------------------------------------------------------------
def makefilter(grid: Grid, order=4, innerK=0.65, outerK=1.0, tol=1e-15) -> np.ndarray:
    """filter(K) = exp(-decay (K - innerK)^order), decay = -log(tol)/(outerK - innerK)^order,
    set to 1 where K < innerK; K = sqrt((kr dx/pi)^2 + (l dy/pi)^2 + (m dz/pi)^2) on the
    half spectrum (realvars)."""
    K2 = (grid.kr * (grid.Lx / grid.nx) / np.pi) ** 2
    if grid.ndim >= 2:
        K2 = K2 + (grid.l * (grid.Ly / grid.ny) / np.pi) ** 2
    if grid.ndim == 3:
        K2 = K2 + (grid.m * (grid.Lz / grid.nz) / np.pi) ** 2
    K = np.broadcast_to(np.sqrt(K2), grid.spec_shape)
    decay = -math.log(tol) / (outerK - innerK) ** order
    filt = np.exp(-decay * (K - innerK) ** order)
    return np.where(K < innerK, 1.0, filt)


def aliased_ranges(nk: int, nkr: int, aliased_fraction: float = 1 / 3):
    """0-based index ranges [lo, hi) of the aliased wavenumbers of a full (fftfreq) axis
    of length nk and of the half (rfftfreq) axis of length nkr:
    iL = floor((1 - f)/2 nk) + 1, iR = ceil((1 + f)/2 nk) as 1-based indices."""
    if not aliased_fraction < 1:
        raise ValueError("aliased_fraction must be less than 1")
    if aliased_fraction <= 0:
        return (nk // 2, nk // 2 + 1), (nkr - 1, nkr)
    iL = math.floor((1 - aliased_fraction) / 2 * nk) + 1
    iR = math.ceil((1 + aliased_fraction) / 2 * nk)
    return (iL - 1, iR), (iL - 1, nkr)


def dealias(grid: Grid, fh: np.ndarray, aliased_fraction: float = 1 / 3) -> np.ndarray:
    """dealias!(fh, grid) on a half-spectrum array (a copy is returned): zero the aliased
    kr columns and the aliased l (and m) rows."""
    out = np.array(fh, copy=True)
    if aliased_fraction <= 0:
        return out
    _, (rlo, rhi) = aliased_ranges(grid.nx, grid.nkr, aliased_fraction)
    out[..., rlo:rhi] = 0
    if grid.ndim >= 2:
        (lo, hi), _ = aliased_ranges(grid.ny, grid.ny // 2 + 1, aliased_fraction)
        out[..., lo:hi, :] = 0
    if grid.ndim == 3:
        (lo, hi), _ = aliased_ranges(grid.nz, grid.nz // 2 + 1, aliased_fraction)
        out[lo:hi, :, :] = 0
    return out


# ----------------------------------------------------------------------------
# Problem / Clock / steppers  ([FF] problem.jl, timesteppers.jl; TAD S:143-216)
# ----------------------------------------------------------------------------
class Clock:
    def __init__(self, dt):
        self.dt = float(dt)
        self.t = 0.0
        self.step = 0


class Problem:
    """Oracle problem: FP64 state + RK4 or ETDRK4 stepping with the reference algebra.

    ``flow`` is either a tuple of steady physical arrays (u[, v[, w]]) or a
    callable ``flow(t) -> tuple`` evaluated once per step at clock.t, the
    time-varying semantics of S:701, S:718, S:737 (the stage time passed to
    calcN! is unused there).
    """

    def __init__(self, grid: Grid, flow, kappa=0.1, eta=None, iota=None, kappah=0.0, nkappah=0,
                 dt=0.01, stepper="RK4", **filterkw):
        self.grid = grid
        self.flow = flow
        self.steady = not callable(flow)
        self.L = linear_operator(grid, kappa, eta, iota, kappah, nkappah)
        self.clock = Clock(dt)
        st = stepper.upper()
        self.filter = None
        if st.startswith("FILTERED"):
            self.filter = makefilter(grid, **filterkw)
            st = st[len("FILTERED"):]
        self.stepper = st
        self.sol = np.zeros(grid.spec_shape, dtype=np.complex128)
        self.c = np.zeros(grid.shape)
        if self.stepper == "ETDRK4":
            self.expLdt, self.expLdt2, self.zeta, self.alpha, self.beta, self.gamma = \
                etd_coefficients(self.clock.dt, self.L)
        elif self.stepper != "RK4":
            raise ValueError(f"unsupported stepper {stepper}")

    # S:844-852
    def set_c(self, c):
        self.sol = rfft(self.grid, c)
        self.updatevars()

    # S:815-821
    def updatevars(self):
        self.c = irfft(self.grid, self.sol.copy())
        return self.c

    def _flow_now(self):
        if self.steady:
            return self.flow
        return self.flow(self.clock.t)

    def _N(self, sol, flow):
        return calcN(self.grid, sol, flow)

    def stepforward(self, nsteps=1):
        for _ in range(int(nsteps)):
            if self.stepper == "RK4":
                self._rk4()
            else:
                self._etdrk4()
            if self.filter is not None:   # [FF] Filtered*update!: sol = filter * update
                self.sol = self.filter * self.sol
            self.clock.t += self.clock.dt
            self.clock.step += 1

    def _rk4(self):
        dt, L, sol = self.clock.dt, self.L, self.sol
        flow = self._flow_now()
        RHS1 = self._N(sol, flow)
        RHS1 += L * sol
        sol1 = sol + (dt / 2) * RHS1
        RHS2 = self._N(sol1, flow)
        RHS2 += L * sol1
        sol1 = sol + (dt / 2) * RHS2
        RHS3 = self._N(sol1, flow)
        RHS3 += L * sol1
        sol1 = sol + dt * RHS3
        RHS4 = self._N(sol1, flow)
        RHS4 += L * sol1
        self.sol = sol + dt * (RHS1 / 6 + RHS2 / 3 + RHS3 / 3 + RHS4 / 6)

    def _etdrk4(self):
        sol = self.sol
        flow = self._flow_now()
        N1 = self._N(sol, flow)
        sol1 = self.expLdt2 * sol + self.zeta * N1
        N2 = self._N(sol1, flow)
        sol2 = self.expLdt2 * sol + self.zeta * N2
        N3 = self._N(sol2, flow)
        sol2 = self.expLdt2 * sol1 + self.zeta * (2 * N3 - N1)
        N4 = self._N(sol2, flow)
        self.sol = (self.expLdt * sol + self.alpha * N1 + 2 * self.beta * (N2 + N3)
                    + self.gamma * N4)


# ----------------------------------------------------------------------------
# diagnostics (FP64)
# ----------------------------------------------------------------------------
def mass(grid: Grid, sol: np.ndarray) -> float:
    """Sum of c over the grid = Re of the k=0 mode of the unnormalised transform."""
    return float(sol.reshape(-1)[0].real)


def variance(grid: Grid, sol: np.ndarray) -> float:
    """(1/N) sum (c - mean c)^2 by Parseval on the half spectrum (weights 1 at kx=0, nx/2; 2 else)."""
    w = np.full(grid.nkr, 2.0)
    w[0] = 1.0
    w[-1] = 1.0
    p = (np.abs(sol) ** 2) * w
    tot = p.sum() - (np.abs(sol.reshape(-1)[0]) ** 2)
    return float(tot / grid.npoints ** 2)
