"""FilteredRK4 / FilteredETDRK4 ([FF] Filtered*TimeStepper: sol = filter * update after
every step, filter = makefilter(grid; order, innerK, outerK, tol)) and dealias! on the
device, against the oracle.

Bars: FP64 relative L2 <= 1e-12 after 10 steps; mixed form <= 1e-5; dealias exact.
"""
import numpy as np
import pytest

import passivetracerflows as ptf
from oracle import tad as O

pytestmark = pytest.mark.gpu
TAD = ptf.TracerAdvectionDiffusion


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def _setup(ndim, n, seed):
    rng = np.random.default_rng(seed)
    g = O.Grid(ndim, n[0]) if ndim == 1 else (O.Grid(2, n[0], ny=n[1]) if ndim == 2
                                             else O.Grid(3, n[0], ny=n[1], nz=n[2]))

    def rough():   # energy up to the grid scale, so that the filter matters
        f = rng.standard_normal(g.shape)
        return O.irfft(g, O.rfft(g, f) * np.exp(-g.Krsq / (4.0 * g.nx)))
    return g, tuple(rough() / 4 for _ in range(ndim)), rough()


def _device(dev, ndim, n, flows, T, stepper, dt, kappa, fk):
    f = [lambda *a, i=i: flows[i] for i in range(ndim)]
    if ndim == 1:
        gr = ptf.OneDGrid(dev, nx=n[0], T=T)
    elif ndim == 2:
        gr = ptf.TwoDGrid(dev, nx=n[0], ny=n[1], T=T)
    else:
        gr = ptf.ThreeDGrid(dev, nx=n[0], ny=n[1], nz=n[2], T=T)
    pr = TAD.ConstDiffSteadyFlowParams(*([kappa] * ndim), *f[:ndim], gr)
    return ptf.FourierFlows.Problem(TAD.Equation(pr, gr), stepper, dt, gr, TAD.Vars(gr), pr, **fk)


CASES = [
    (1, (128,), "FilteredRK4", {}), (2, (128, 64), "FilteredRK4", {}),
    (2, (64, 128), "FilteredETDRK4", dict(order=6, innerK=0.5, outerK=0.9, tol=1e-12)),
    (3, (32, 32, 64), "FilteredRK4", dict(order=4, innerK=0.6, outerK=1.0, tol=1e-15)),
    (3, (32, 16, 32), "FilteredETDRK4", {}),
]


@pytest.mark.parametrize("ndim,n,stepper,fk", CASES)
@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_filtered_stepper_parity(dev, ndim, n, stepper, fk, T):
    print(f"filtered {stepper} ndim={ndim} n={n} {T.__name__}", flush=True)
    g, flows, c0 = _setup(ndim, n, seed=5 + ndim)
    dt, kappa, nsteps = 0.01, 1e-3, 10
    o = O.Problem(g, flows, kappa=kappa, eta=kappa if ndim >= 2 else None,
                  iota=kappa if ndim == 3 else None, dt=dt, stepper=stepper, **fk)
    o.set_c(c0)
    o.stepforward(nsteps)
    p = _device(dev, ndim, n, flows, T, stepper, dt, kappa, fk)
    TAD.set_c_(p, c0)
    ptf.stepforward_(p, nsteps)
    # the filter acts: the unfiltered run differs from the filtered one
    u = O.Problem(g, flows, kappa=kappa, eta=kappa if ndim >= 2 else None,
                  iota=kappa if ndim == 3 else None, dt=dt, stepper=stepper[len("Filtered"):])
    u.set_c(c0)
    u.stepforward(nsteps)
    assert rel(u.sol, o.sol) > 1e-6
    assert rel(p.sol, o.sol) <= (1e-12 if T == np.float64 else 1e-5)


def test_tad_problem_filtered_default(dev):
    """TracerAdvectionDiffusion.Problem(...; stepper="FilteredRK4") takes FourierFlows'
    default filter (order 4, innerK 0.65, outerK 1, tol 1e-15)."""
    g, flows, c0 = _setup(2, (64, 64), seed=9)
    p = TAD.Problem(dev, ptf.TwoDAdvectingFlow(u=lambda x, y: flows[0], v=lambda x, y: flows[1]),
                    nx=64, κ=1e-3, dt=0.01, stepper="FilteredRK4")
    TAD.set_c_(p, c0)
    ptf.stepforward_(p, 10)
    o = O.Problem(g, flows, kappa=1e-3, dt=0.01, stepper="FilteredRK4")
    o.set_c(c0)
    o.stepforward(10)
    assert rel(p.sol, o.sol) <= 1e-12
    with pytest.raises(TypeError):
        ptf.FourierFlows.Problem(p.eqn, "RK4", 0.01, p.grid, p.vars, p.params, order=4)


@pytest.mark.parametrize("ndim,n", [(1, (64,)), (2, (32, 64)), (3, (16, 32, 8))])
@pytest.mark.parametrize("T", [np.float64, np.float32])
@pytest.mark.parametrize("frac", [1 / 3, 0.5, 0.0])
def test_dealias_device(dev, ndim, n, T, frac):
    g = O.Grid(ndim, n[0]) if ndim == 1 else (O.Grid(2, n[0], ny=n[1]) if ndim == 2
                                             else O.Grid(3, n[0], ny=n[1], nz=n[2]))
    kw = dict(nx=n[0], T=T, aliased_fraction=frac)
    gr = (ptf.OneDGrid(dev, **kw) if ndim == 1 else
          ptf.TwoDGrid(dev, ny=n[1], **kw) if ndim == 2 else ptf.ThreeDGrid(dev, ny=n[1], nz=n[2], **kw))
    rng = np.random.default_rng(1)
    fh = (rng.standard_normal(g.spec_shape) + 1j * rng.standard_normal(g.spec_shape))
    ct = np.complex128 if T == np.float64 else np.complex64
    d = ptf.DeviceArray.from_host(fh.astype(ct))
    ptf.dealias_(d, gr)
    assert np.array_equal(d.to_host(), O.dealias(g, fh, frac).astype(ct))
