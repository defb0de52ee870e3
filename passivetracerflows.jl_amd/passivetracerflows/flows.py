"""Built-in analytic flows evaluated on the GPU (no per-step host evaluation or upload).

A flow is a list of separable terms; the velocity component ``c`` is

    sum_k  amp_k * T_k(omega_k t + tphase_k) * X_k(kx_k x + px_k) * Y_k(...) * Z_k(...)

with X, Y, Z, T each one of ``"1"``, ``"sin"``, ``"cos"`` (``ptf_flow_term`` in
include/ptf.h; the GPU multiplies ((amp T) X) Y) Z and sums the terms of a
component in order).  The x/y/z factors are tabulated once on the grid points,
the time factors at clock.t before every step, so a time-varying flow of this
form costs one flow-synthesis kernel per step instead of the reference's
host-side ``u(x, y, t)`` evaluation (S:701, S:718, S:737) plus a PCIe upload.

Constructors of common flows: ``cellular_flow`` (psi = (U/k) sin kx sin ky),
``abc_flow`` (Arnold-Beltrami-Childress), ``shear_flow``, ``uniform_flow``;
each takes an optional time modulation.
"""
from __future__ import annotations

from dataclasses import dataclass

from .traceradvectiondiffusion import AbstractAdvectingFlow

_FN = {"1": 0, "sin": 1, "cos": 2}


@dataclass
class FlowTerm:
    component: int                  # 0 = u, 1 = v, 2 = w
    amp: float = 1.0
    fx: str = "1"
    fy: str = "1"
    fz: str = "1"
    kx: float = 0.0
    ky: float = 0.0
    kz: float = 0.0
    px: float = 0.0
    py: float = 0.0
    pz: float = 0.0
    time: str = "1"
    omega: float = 0.0
    tphase: float = 0.0

    def ctype(self):
        from . import _lib
        t = _lib.ptf_flow_term()
        t.component = int(self.component)
        t.fx, t.fy, t.fz, t.time_fn = (_FN[self.fx], _FN[self.fy], _FN[self.fz], _FN[self.time])
        t.kx, t.ky, t.kz = float(self.kx), float(self.ky), float(self.kz)
        t.px, t.py, t.pz = float(self.px), float(self.py), float(self.pz)
        t.amp, t.omega, t.tphase = float(self.amp), float(self.omega), float(self.tphase)
        return t


class AnalyticAdvectingFlow(AbstractAdvectingFlow):
    """An advecting flow given as FlowTerms, synthesised on the GPU; ``ndim`` is the
    dimension of the problem it advects in (1, 2 or 3)."""

    def __init__(self, ndim: int, terms):
        self.ndim = int(ndim)
        self.terms = list(terms)
        if not 1 <= len(self.terms) <= 16:
            raise ValueError("an analytic flow has 1 to 16 terms")
        for t in self.terms:
            if not 0 <= t.component < self.ndim:
                raise ValueError("term component out of range for the flow's dimension")
            for f in (t.fx, t.fy, t.fz, t.time):
                if f not in _FN:
                    raise ValueError(f"unknown factor {f!r} (use '1', 'sin' or 'cos')")
        self.steadyflow = all(t.time == "1" for t in self.terms)


def _modulate(terms, time, omega, tphase):
    for t in terms:
        t.time, t.omega, t.tphase = time, omega, tphase
    return terms


def cellular_flow(U=1.0, k=1.0, time="1", omega=0.0, tphase=0.0):
    """2D cellular flow from psi = (U/k) sin(kx) sin(ky): u = -d_y psi = -U sin(kx) cos(ky),
    v = d_x psi = U cos(kx) sin(ky); amplitude times T(omega t + tphase)."""
    terms = [FlowTerm(0, -U, fx="sin", fy="cos", kx=k, ky=k),
             FlowTerm(1, U, fx="cos", fy="sin", kx=k, ky=k)]
    return AnalyticAdvectingFlow(2, _modulate(terms, time, omega, tphase))


def abc_flow(A=1.0, B=1.0, C=1.0, time="1", omega=0.0, tphase=0.0):
    """3D ABC flow u = A sin z + C cos y, v = B sin x + A cos z, w = C sin y + B cos x."""
    terms = [FlowTerm(0, A, fz="sin", kz=1.0), FlowTerm(0, C, fy="cos", ky=1.0),
             FlowTerm(1, B, fx="sin", kx=1.0), FlowTerm(1, A, fz="cos", kz=1.0),
             FlowTerm(2, C, fy="sin", ky=1.0), FlowTerm(2, B, fx="cos", kx=1.0)]
    return AnalyticAdvectingFlow(3, _modulate(terms, time, omega, tphase))


def shear_flow(U=1.0, k=1.0, ndim=2, time="1", omega=0.0, tphase=0.0):
    """Parallel shear u = U sin(k y)."""
    return AnalyticAdvectingFlow(ndim, _modulate([FlowTerm(0, U, fy="sin", ky=k)], time, omega,
                                                 tphase))


def uniform_flow(*U, time="1", omega=0.0, tphase=0.0):
    """Uniform flow (U[, V[, W]]); its dimension is the number of components."""
    terms = [FlowTerm(c, float(u)) for c, u in enumerate(U)]
    return AnalyticAdvectingFlow(len(U), _modulate(terms, time, omega, tphase))
