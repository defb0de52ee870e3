"""Graph replay of steady-flow steps (ptf_step captures 8-step blocks) against direct
launches: every replay, with eager work and synchronisation between replays, gives
bitwise the same state; the clock advances as with direct steps; a variant change
re-captures."""
import numpy as np
import pytest

import passivetracerflows as ptf
from oracle import tad as O

pytestmark = pytest.mark.gpu
TAD = ptf.TracerAdvectionDiffusion


def _setup(ndim, n, seed):
    rng = np.random.default_rng(seed)
    g = O.Grid(2, n[0], ny=n[1]) if ndim == 2 else O.Grid(3, n[0], ny=n[1], nz=n[2])

    def smooth():
        return O.irfft(g, O.rfft(g, rng.standard_normal(g.shape)) * np.exp(-g.Krsq / 8.0))
    return g, tuple(smooth() / 2 for _ in range(ndim)), smooth()


def _problem(dev, ndim, n, flows, T, stepper):
    f = [lambda *a, i=i: flows[i] for i in range(ndim)]
    if ndim == 2:
        return TAD.Problem(dev, ptf.TwoDAdvectingFlow(u=f[0], v=f[1]), nx=n[0], ny=n[1], κ=1e-3,
                           dt=0.01, stepper=stepper, T=T)
    return TAD.Problem(dev, ptf.ThreeDAdvectingFlow(u=f[0], v=f[1], w=f[2]), nx=n[0], ny=n[1],
                       nz=n[2], κ=1e-3, dt=0.01, stepper=stepper, T=T)


@pytest.mark.parametrize("ndim,n,stepper,T", [
    (3, (64, 64, 64), "RK4", np.float64), (2, (128, 64), "ETDRK4", np.float32),
    (2, (64, 128), "FilteredRK4", np.float64)])
def test_graph_replay_equals_direct(dev, ndim, n, stepper, T):
    print(f"graph replay ndim={ndim} n={n} {stepper} {T.__name__}", flush=True)
    g, flows, c0 = _setup(ndim, n, 21)
    a = _problem(dev, ndim, n, flows, T, stepper)
    b = _problem(dev, ndim, n, flows, T, stepper)
    b.set_graph(False)
    for p in (a, b):
        TAD.set_c_(p, c0)
    for nsteps in (16, 13, 8, 24):          # whole blocks and remainders
        ptf.stepforward_(a, nsteps)
        ptf.stepforward_(b, nsteps)
        assert np.array_equal(a.sol, b.sol), nsteps
        assert a.clock.step == b.clock.step and a.clock.t == b.clock.t
        # eager work between replays
        TAD.updatevars_(a)
        assert TAD.mass(a) == TAD.mass(b)
    # a variant change drops the captured graph; the next steps re-capture
    a.autotune()
    ptf.stepforward_(a, 16)
    b.set_graph(True)
    ptf.stepforward_(b, 16)
    o = O.Problem(g, flows, kappa=1e-3, dt=0.01, stepper=stepper)
    o.set_c(c0)
    o.stepforward(a.clock.step)
    bar = 1e-12 if T == np.float64 else 1e-5
    assert float(np.linalg.norm(a.sol - o.sol) / np.linalg.norm(o.sol)) <= bar
    assert float(np.linalg.norm(b.sol - o.sol) / np.linalg.norm(o.sol)) <= bar
