"""Grid sizes that are not powers of two (2^a m, m in {3, 5, 9, 25, 27, 125}).

The reference builds its grids over FFTW plans, which take any even n
(src/traceradvectiondiffusion.jl:152, 177, 205).  Here the same line transforms run
with mixed radix plans (radices 2, 3, 4, 5, 8, 16).  Checked against pocketfft
(the rfft plan, forward and non-Hermitian inverse) and against the FP64 oracle
(whole RK4 and ETDRK4 steps, 1D/2D/3D, bricked and plain layouts, FP64 and mixed).
Bars as for the power-of-two sizes: transforms 1e-13 (FP64) / 2e-6 (FP32); steps
1e-12 (FP64) / 1e-5 (mixed).
"""
import numpy as np
import pytest

import passivetracerflows as ptf
from oracle import tad as O

from test_gpu_step import _device_problem, _rand_setup

pytestmark = pytest.mark.gpu
TAD = ptf.TracerAdvectionDiffusion


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def _grid(dev, ndim, n, T=np.float64):
    if ndim == 1:
        return ptf.OneDGrid(dev, nx=n[0], T=T)
    if ndim == 2:
        return ptf.TwoDGrid(dev, nx=n[0], ny=n[1], T=T)
    return ptf.ThreeDGrid(dev, nx=n[0], ny=n[1], nz=n[2], T=T)


SHAPES = [(1, (96,)), (1, (640,)), (1, (1000,)), (1, (4000,)), (1, (54,)),
          (2, (96, 192)), (2, (384, 96)), (2, (1000, 64)), (2, (64, 1000)), (2, (640, 640)),
          (2, (48, 250)), (3, (96, 96, 96)), (3, (40, 36, 192)), (3, (200, 24, 1000))]


@pytest.mark.parametrize("ndim,n", SHAPES)
@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_rfft_irfft_sizes_vs_pocketfft(dev, ndim, n, T):
    print(f"sizes rfft/irfft ndim={ndim} n={n} T={T.__name__}", flush=True)
    g = _grid(dev, ndim, n, T)
    rng = np.random.default_rng(5)
    c = rng.standard_normal(g.shape)
    ct = np.complex128 if T is np.float64 else np.complex64
    tol = 1e-13 if T is np.float64 else 2e-6
    din = ptf.DeviceArray.from_host(c.astype(T))
    dout = ptf.DeviceArray(g.spec_shape, ct)
    ptf.mul_(dout, g.rfftplan, din)
    assert rel(dout.to_host(), np.fft.rfftn(c, axes=tuple(range(ndim)))) < tol
    sp = rng.standard_normal(g.spec_shape) + 1j * rng.standard_normal(g.spec_shape)
    dsp = ptf.DeviceArray.from_host(sp.astype(ct))
    dphys = ptf.DeviceArray(g.shape, T)
    ptf.ldiv_(dphys, g.rfftplan, dsp)
    refb = np.fft.irfftn(sp, s=g.shape, axes=tuple(range(ndim)))
    assert rel(dphys.to_host(), refb) < tol
    assert np.array_equal(dsp.to_host(), sp.astype(ct))


STEP_CASES = [
    (1, (96,), "RK4", 0.0, 0), (1, (1000,), "ETDRK4", 1e-9, 2),
    (2, (96, 192), "RK4", 0.0, 0), (2, (384, 96), "ETDRK4", 1e-8, 2),
    (2, (1000, 64), "RK4", 0.0, 0), (2, (64, 640), "RK4", 1e-9, 2),
    (3, (96, 48, 96), "RK4", 0.0, 0), (3, (40, 36, 192), "ETDRK4", 1e-6, 2),
]


@pytest.mark.parametrize("ndim,n,stepper,kappah,nkappah", STEP_CASES)
@pytest.mark.parametrize("T,bar", [(np.float64, 1e-12), (np.float32, 1e-5)])
def test_step_sizes_vs_oracle(dev, ndim, n, stepper, kappah, nkappah, T, bar):
    print(f"sizes step ndim={ndim} n={n} {stepper} {T.__name__}", flush=True)
    g, flows, c0 = _rand_setup(ndim, n, seed=17 + ndim)
    dt, nsteps = 0.005, 8
    kappa = 1e-3 if max(n) >= 512 else 1e-2
    eta = kappa / 2 if ndim >= 2 else None
    iota = kappa / 3 if ndim == 3 else None
    o = O.Problem(g, flows, kappa=kappa, eta=eta, iota=iota, kappah=kappah, nkappah=nkappah,
                  dt=dt, stepper=stepper)
    o.set_c(c0)
    o.stepforward(nsteps)
    p = _device_problem(dev, ndim, n, flows, T, stepper, dt, kappa, kappah, nkappah)
    TAD.set_c_(p, c0)
    ptf.stepforward_(p, nsteps)
    assert rel(p.sol, o.sol) <= bar
    if T is np.float64:
        TAD.updatevars_(p)
        assert rel(p.vars.c, o.updatevars()) <= bar
    # the autotuned variants agree (steady flow: the clock does not enter the step)
    p.autotune()
    TAD.set_c_(p, c0)
    ptf.stepforward_(p, nsteps)
    assert rel(p.sol, o.sol) <= bar
