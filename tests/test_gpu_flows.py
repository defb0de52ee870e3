"""Built-in analytic flows (passivetracerflows.flows, PTF_FLOW_ANALYTIC) synthesised on
the GPU, against the oracle driven by the same flow as a user function u(x, y, t)
evaluated at clock.t (the reference's time-varying semantics, S:701, S:718, S:737).

Bars: FP64 relative L2 <= 1e-12; mixed form <= 1e-5; a steady analytic flow equals
the same flow sampled into arrays to <= 1e-13.
"""
import numpy as np
import pytest

import passivetracerflows as ptf
from oracle import tad as O
from passivetracerflows import flows as F

pytestmark = pytest.mark.gpu
TAD = ptf.TracerAdvectionDiffusion
_FN = {"1": lambda a: np.ones_like(a), "sin": np.sin, "cos": np.cos}


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def evaluate(flow, g, t):
    """The flow as a user function would compute it: numpy on the grid points."""
    pts = g.gridpoints()
    out = [np.zeros(g.shape) for _ in range(flow.ndim)]
    for term in flow.terms:
        tf = {"1": 1.0, "sin": np.sin(term.omega * t + term.tphase),
              "cos": np.cos(term.omega * t + term.tphase)}[term.time]
        v = term.amp * tf * _FN[term.fx](term.kx * pts[0] + term.px)
        if flow.ndim >= 2:
            v = v * _FN[term.fy](term.ky * pts[1] + term.py)
        if flow.ndim == 3:
            v = v * _FN[term.fz](term.kz * pts[2] + term.pz)
        out[term.component] = out[term.component] + v
    return tuple(out)


def _c0(g, seed):
    rng = np.random.default_rng(seed)
    return O.irfft(g, O.rfft(g, rng.standard_normal(g.shape)) * np.exp(-g.Krsq / 8.0))


CASES = [
    ("cellular_cos", lambda: F.cellular_flow(U=1.3, k=2.0, time="cos", omega=2.0), (64, 64), "RK4"),
    ("abc_steady", lambda: F.abc_flow(1.0, 0.8, 0.6), (32, 32, 32), "RK4"),
    ("abc_sin", lambda: F.abc_flow(1.0, 0.8, 0.6, time="sin", omega=1.5, tphase=0.3),
     (32, 16, 64), "ETDRK4"),
    ("shear_cos", lambda: F.shear_flow(0.7, 3.0, time="cos", omega=4.0), (64, 128), "RK4"),
    ("uniform_1d_sin", lambda: F.uniform_flow(0.4, time="sin", omega=3.0, tphase=0.5), (128,), "RK4"),
]


@pytest.mark.parametrize("name,make,n,stepper", CASES)
@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_analytic_flow_parity(dev, name, make, n, stepper, T):
    print(f"analytic flow {name} n={n} {stepper} {T.__name__}", flush=True)
    flow = make()
    nd = flow.ndim
    g = O.Grid(nd, n[0]) if nd == 1 else (O.Grid(2, n[0], ny=n[1]) if nd == 2
                                         else O.Grid(3, n[0], ny=n[1], nz=n[2]))
    dt, kappa, nsteps = 0.01, 2e-3, 20
    c0 = _c0(g, 3)
    of = (lambda t: evaluate(flow, g, t)) if not flow.steadyflow else evaluate(flow, g, 0.0)
    o = O.Problem(g, of, kappa=kappa, dt=dt, stepper=stepper)
    o.set_c(c0)
    o.stepforward(nsteps)
    kw = dict(nx=n[0], κ=kappa, dt=dt, stepper=stepper, T=T)
    if nd >= 2:
        kw["ny"] = n[1]
    if nd == 3:
        kw["nz"] = n[2]
    p = TAD.Problem(dev, flow, **kw)
    TAD.set_c_(p, c0)
    ptf.stepforward_(p, nsteps)
    assert p.clock.step == nsteps
    assert rel(p.sol, o.sol) <= (1e-12 if T == np.float64 else 1e-5)


def test_steady_analytic_equals_sampled_arrays(dev):
    g = O.Grid(3, 64)
    flow = F.abc_flow(1.0, 0.8, 0.6)
    u, v, w = evaluate(flow, g, 0.0)
    c0 = _c0(g, 4)
    a = TAD.Problem(dev, flow, nx=64, κ=1e-3, dt=5e-3)
    b = TAD.Problem(dev, ptf.ThreeDAdvectingFlow(u=lambda *x: u, v=lambda *x: v, w=lambda *x: w),
                    nx=64, κ=1e-3, dt=5e-3)
    for p in (a, b):
        TAD.set_c_(p, c0)
        ptf.stepforward_(p, 10)
    assert rel(a.sol, b.sol) <= 1e-13


def test_analytic_flow_validation(dev):
    with pytest.raises(ValueError):
        F.AnalyticAdvectingFlow(2, [F.FlowTerm(2, 1.0)])      # w in a 2D flow
    with pytest.raises(ValueError):
        F.AnalyticAdvectingFlow(2, [F.FlowTerm(0, 1.0, fx="tan")])
    with pytest.raises(ValueError):       # more than 16 terms
        F.AnalyticAdvectingFlow(2, [F.FlowTerm(0, 1.0)] * 17)
