"""Every compiled kernel variant of every launch slot against the oracle.

The autotuner may select any (radix, width, register budget, advection form,
bricked-line lane mapping) variant of a slot, so each one is checked: one slot at a time is switched to
the variant (the other slots keep their defaults), the problem restarts from
the same state and takes 3 steps.  Bars: FP64 relative L2 <= 1e-12 against the
FP64 oracle; mixed form <= 1e-5.
"""
import numpy as np
import pytest

import passivetracerflows as ptf
from oracle import tad as O
from passivetracerflows._lib import PTFError

pytestmark = pytest.mark.gpu
TAD = ptf.TracerAdvectionDiffusion

RADICES = (4, 8, 16)
WIDTHS = (1, 2, 4, 8, 16, 32, 64)
FORMS = ((False, False), (True, False), (False, True), (True, True))   # (occ3, par)


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def _setup(ndim, n, seed):
    rng = np.random.default_rng(seed)
    g = O.Grid(ndim, n[0]) if ndim == 1 else (O.Grid(2, n[0], ny=n[1]) if ndim == 2
                                             else O.Grid(3, n[0], ny=n[1], nz=n[2]))

    def smooth():
        f = rng.standard_normal(g.shape)
        return O.irfft(g, O.rfft(g, f) * np.exp(-g.Krsq / 8.0))
    return g, tuple(smooth() / 2 for _ in range(ndim)), smooth()


def _problem(dev, ndim, n, flows, T, stepper, dt, kappa):
    f = [lambda *a, i=i: flows[i] for i in range(ndim)]
    if ndim == 1:
        return TAD.Problem(dev, ptf.OneDAdvectingFlow(u=f[0]), nx=n[0], κ=kappa, dt=dt,
                           stepper=stepper, T=T)
    if ndim == 2:
        return TAD.Problem(dev, ptf.TwoDAdvectingFlow(u=f[0], v=f[1]), nx=n[0], ny=n[1], κ=kappa,
                           dt=dt, stepper=stepper, T=T)
    return TAD.Problem(dev, ptf.ThreeDAdvectingFlow(u=f[0], v=f[1], w=f[2]), nx=n[0], ny=n[1],
                       nz=n[2], κ=kappa, dt=dt, stepper=stepper, T=T)


CASES = [
    (3, (64, 64, 64), "RK4", np.float64), (3, (64, 64, 64), "RK4", np.float32),
    (3, (32, 32, 64), "ETDRK4", np.float64),
    (2, (128, 128), "RK4", np.float64), (2, (256, 64), "ETDRK4", np.float32),
    (1, (256,), "RK4", np.float64),
    # long lines (>= 512): one-line-per-block column variants and part-brick row blocks
    (2, (64, 512), "RK4", np.float64), (2, (128, 512), "ETDRK4", np.float32),
    (3, (16, 512, 64), "RK4", np.float32), (3, (16, 16, 512), "ETDRK4", np.float64),
    # 1024- and 2048-point lines: twiddle tables in LDS at their largest, and the
    # (width, radix) variants that still fit the 160 KiB LDS budget
    (2, (2048, 64), "RK4", np.float64), (2, (64, 2048), "ETDRK4", np.float32),
    (3, (16, 16, 1024), "RK4", np.float64), (2, (1024, 64), "ETDRK4", np.float32),
    # FP64 2048-point columns keep half of their twiddle table in LDS (w[k + N/2] = -w[k])
    (2, (64, 2048), "ETDRK4", np.float64), (2, (64, 2048), "RK4", np.float64),
]


@pytest.mark.parametrize("ndim,n,stepper,T", CASES)
def test_every_variant_of_every_slot(dev, ndim, n, stepper, T):
    print(f"variants ndim={ndim} n={n} {stepper} {T.__name__}", flush=True)
    g, flows, c0 = _setup(ndim, n, seed=11 + ndim)
    # kappa |k_max|^2 dt stays inside RK4's real-axis stability limit for lines of 512..2048
    kappa = 1e-4 if max(n) >= 1024 else (1e-3 if max(n) >= 512 else 1e-2)
    dt, nsteps = 0.01, 3
    o = O.Problem(g, flows, kappa=kappa, dt=dt, stepper=stepper)
    o.set_c(c0)
    sol0 = o.sol.copy()
    o.stepforward(nsteps)
    bar = 1e-12 if T == np.float64 else 1e-5
    p = _problem(dev, ndim, n, flows, T, stepper, dt, kappa)
    names = p.slots()
    tested = 0
    for i, name in enumerate(names):
        default = p.get_variant(i)
        count = 0
        for R in RADICES:
            for W in WIDTHS:
                for occ3, par in FORMS:
                    for wg in (False, True):
                        try:
                            p.set_variant(i, R, W, occ3=occ3, par=par, wg=wg)
                        except PTFError:
                            continue
                        p.sol = sol0
                        ptf.stepforward_(p, nsteps)
                        err = rel(p.sol, o.sol)
                        assert err <= bar, (name, R, W, occ3, par, wg, err)
                        count += 1
        assert count >= 1, name
        tested += count
        R, W, occ3, par, wg = default
        p.set_variant(i, R, W, occ3=occ3, par=par, wg=wg)
    print(f"  {len(names)} slots, {tested} variants", flush=True)
    with pytest.raises(PTFError):
        p.set_variant(0, 3, 8)
