"""The alternative launch schedules and lane mappings against the oracle.

The default schedule fuses each stage's last forward pass with the next stage's first
inverse pass (PTF_STAGE_FUSION=0, read when a problem is created, builds the unfused
schedule: each stage starts with its own inverse pass over the stored stage input).  By
default a column line along the bricked axis is held by one wavefront where its threads
fit; the WG_LINES variant flag spreads it over the workgroup with workgroup barriers.
Every combination, with every slot that accepts the flag switched to it: FP64 relative
L2 <= 1e-12 against the FP64 oracle after 3 steps; mixed form <= 1e-5.
"""
import numpy as np
import pytest

import passivetracerflows as ptf
from oracle import tad as O
from passivetracerflows._lib import PTFError

pytestmark = pytest.mark.gpu
TAD = ptf.TracerAdvectionDiffusion


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def _case(ndim, n, seed):
    rng = np.random.default_rng(seed)
    g = O.Grid(2, n[0], ny=n[1]) if ndim == 2 else O.Grid(3, n[0], ny=n[1], nz=n[2])

    def smooth():
        f = rng.standard_normal(g.shape)
        return O.irfft(g, O.rfft(g, f) * np.exp(-g.Krsq / 8.0))
    return g, tuple(smooth() / 2 for _ in range(ndim)), smooth()


def _workgroup_lines(p):
    """Switch every slot that has the per-wavefront mapping to the workgroup one."""
    switched = 0
    for i in range(len(p.slots())):
        R, W, occ3, par, wg = p.get_variant(i)
        try:
            p.set_variant(i, R, W, occ3=occ3, par=par, wg=True)
            switched += 1
        except PTFError:
            pass
    return switched


CASES = [(3, (64, 64, 64), "RK4"), (2, (128, 256), "ETDRK4"), (2, (96, 192), "RK4")]


@pytest.mark.parametrize("ndim,n,stepper", CASES)
def test_schedules_vs_oracle(dev, monkeypatch, ndim, n, stepper):
    g, flows, c0 = _case(ndim, n, 11)
    dt, kappa = 2e-3, 1e-2
    o = O.Problem(g, flows, kappa=kappa, dt=dt, stepper=stepper)
    o.set_c(c0)
    o.stepforward(3)
    f = [lambda *a, i=i: flows[i] for i in range(ndim)]
    for fusion in ("1", "0"):
        monkeypatch.setenv("PTF_STAGE_FUSION", fusion)
        for wg in (False, True):
            for T, tol in ((np.float64, 1e-12), (np.float32, 1e-5)):
                print(f"{ndim}D {n} {stepper} {T.__name__} fusion={fusion} wg={wg}", flush=True)
                if ndim == 2:
                    p = TAD.Problem(dev, ptf.TwoDAdvectingFlow(u=f[0], v=f[1]), nx=n[0], ny=n[1],
                                    κ=kappa, dt=dt, stepper=stepper, T=T)
                else:
                    p = TAD.Problem(dev, ptf.ThreeDAdvectingFlow(u=f[0], v=f[1], w=f[2]), nx=n[0],
                                    ny=n[1], nz=n[2], κ=kappa, dt=dt, stepper=stepper, T=T)
                if wg:
                    assert _workgroup_lines(p) >= 1
                TAD.set_c_(p, c0)
                ptf.stepforward_(p, 3)
                assert rel(p.sol, o.sol) <= tol, (fusion, wg, T)
