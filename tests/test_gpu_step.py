"""GPU time-step parity: the reference's analytic tests, and device vs oracle.

Bars (written per test): FP64 relative L2 <= 1e-12 against the oracle after
the stated steps, mass conserved to roundoff; the mixed FP32 form <= 1e-5
against the FP64 oracle on these short runs.
"""
import numpy as np
import pytest

import passivetracerflows as ptf
from oracle import tad as O
from passivetracerflows import workloads as W
from reference_cases import NAMES, case, isapprox, run_device, run_oracle

pytestmark = pytest.mark.gpu
TAD = ptf.TracerAdvectionDiffusion


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


@pytest.mark.parametrize("name", NAMES)
def test_reference_analytic_case(dev, name):
    """test/runtests.jl:26-54 on the device, at the reference tolerances, plus parity
    against the oracle (relL2 <= 1e-12)."""
    print(f"reference case {name}", flush=True)
    c = case(name)
    got, prob = run_device(dev, c)
    assert isapprox(c["cfinal"], got, c["rtol"]), name
    assert prob.clock.step == c["nsteps"]
    assert abs(prob.clock.t - c["nsteps"] * c["dt"]) < 1e-12
    ref = run_oracle(c)
    assert rel(got, ref) <= 1e-12, name


def test_noflow():
    assert TAD.noflow(np.pi) == 0


def _rand_setup(ndim, n, seed):
    rng = np.random.default_rng(seed)
    g = O.Grid(ndim, *n) if ndim == 1 else (O.Grid(2, n[0], ny=n[1]) if ndim == 2
                                           else O.Grid(3, n[0], ny=n[1], nz=n[2]))
    # smooth random flow and tracer: low-pass random fields
    def smooth():
        f = rng.standard_normal(g.shape)
        fh = O.rfft(g, f) * np.exp(-g.Krsq / 8.0)
        return O.irfft(g, fh)
    flows = tuple(smooth() / 2 for _ in range(ndim))
    c0 = smooth()
    return g, flows, c0


def _device_problem(dev, ndim, n, flows, T, stepper, dt, kappa, kappah=0.0, nkappah=0):
    f = [lambda *a, i=i: flows[i] for i in range(ndim)]
    if ndim == 1:
        if kappah:
            gr = ptf.OneDGrid(dev, nx=n[0], T=T)
            pr = TAD.ConstDiffSteadyFlowParams(kappa, kappah, nkappah, f[0], gr)
            return ptf.FourierFlows.Problem(TAD.Equation(pr, gr), stepper, dt, gr, TAD.Vars(gr), pr)
        return TAD.Problem(dev, ptf.OneDAdvectingFlow(u=f[0]), nx=n[0], κ=kappa, dt=dt,
                           stepper=stepper, T=T)
    if ndim == 2:
        return TAD.Problem(dev, ptf.TwoDAdvectingFlow(u=f[0], v=f[1]), nx=n[0], ny=n[1], κ=kappa,
                           η=kappa / 2, κh=kappah, nκh=nkappah, dt=dt, stepper=stepper, T=T)
    if kappah:
        gr = ptf.ThreeDGrid(dev, nx=n[0], ny=n[1], nz=n[2], T=T)
        pr = TAD.ConstDiffSteadyFlowParams(kappa, kappa / 2, kappa / 3, kappah, nkappah,
                                           f[0], f[1], f[2], gr)
        return ptf.FourierFlows.Problem(TAD.Equation(pr, gr), stepper, dt, gr, TAD.Vars(gr), pr)
    return TAD.Problem(dev, ptf.ThreeDAdvectingFlow(u=f[0], v=f[1], w=f[2]), nx=n[0], ny=n[1],
                       nz=n[2], κ=kappa, η=kappa / 2, ι=kappa / 3, dt=dt, stepper=stepper, T=T)


CASES = [
    (1, (64,), "RK4", 0.0, 0), (1, (256,), "ETDRK4", 1e-6, 2),
    (2, (64, 32), "RK4", 0.0, 0), (2, (128, 128), "ETDRK4", 1e-8, 3),
    (2, (256, 64), "RK4", 1e-7, 2),
    (3, (32, 16, 64), "RK4", 0.0, 0), (3, (32, 32, 32), "ETDRK4", 1e-6, 2),
]


@pytest.mark.parametrize("ndim,n,stepper,kappah,nkappah", CASES)
def test_step_parity_fp64(dev, ndim, n, stepper, kappah, nkappah):
    print(f"step parity fp64 ndim={ndim} n={n} {stepper}", flush=True)
    g, flows, c0 = _rand_setup(ndim, n, seed=ndim * 7 + len(stepper))
    dt, kappa, nsteps = 0.01, 1e-2, 12
    eta = kappa / 2 if ndim >= 2 else None
    iota = kappa / 3 if ndim == 3 else None
    o = O.Problem(g, flows, kappa=kappa, eta=eta, iota=iota, kappah=kappah, nkappah=nkappah,
                  dt=dt, stepper=stepper)
    o.set_c(c0)
    o.stepforward(nsteps)
    p = _device_problem(dev, ndim, n, flows, np.float64, stepper, dt, kappa, kappah, nkappah)
    TAD.set_c_(p, c0)
    m0 = TAD.mass(p)
    ptf.stepforward_(p, nsteps)
    TAD.updatevars_(p)
    assert rel(p.sol, o.sol) <= 1e-12
    assert rel(p.vars.c, o.updatevars()) <= 1e-12
    # mass: the k = 0 mode is only changed by Re N(0), which vanishes to roundoff
    # for these spectrally smooth flows
    assert abs(TAD.mass(p) - O.mass(g, o.sol)) <= 1e-12 * abs(np.abs(o.sol).max())
    assert abs(TAD.variance(p) - O.variance(g, o.sol)) <= 1e-12 * O.variance(g, o.sol)
    assert p.clock.step == nsteps


@pytest.mark.parametrize("ndim,n,stepper,kappah,nkappah", CASES)
def test_step_mixed_vs_fp64_oracle(dev, ndim, n, stepper, kappah, nkappah):
    print(f"step mixed ndim={ndim} n={n} {stepper}", flush=True)
    g, flows, c0 = _rand_setup(ndim, n, seed=ndim * 7 + len(stepper))
    dt, kappa, nsteps = 0.01, 1e-2, 12
    eta = kappa / 2 if ndim >= 2 else None
    iota = kappa / 3 if ndim == 3 else None
    o = O.Problem(g, flows, kappa=kappa, eta=eta, iota=iota, kappah=kappah, nkappah=nkappah,
                  dt=dt, stepper=stepper)
    o.set_c(c0)
    o.stepforward(nsteps)
    p = _device_problem(dev, ndim, n, flows, np.float32, stepper, dt, kappa, kappah, nkappah)
    TAD.set_c_(p, c0)
    ptf.stepforward_(p, nsteps)
    assert rel(p.sol, o.sol) <= 1e-5


def test_spectral_flow_etdrk4_parity(dev):
    """Built-in random-phase streamfunction flow (small turbulent_2d), ETDRK4 with
    hyperdiffusion order 4, flow evaluated at clock.t each step."""
    n, nsteps, dt = 64, 20, 2e-3
    psih0, omega = W.turbulent_psih(n, seed=4)
    g = O.Grid(2, n)
    X, Y = W.grid_xyz(2, n)
    rng = np.random.default_rng(104)
    c0 = np.sin(X) + 0.01 * rng.standard_normal(X.shape)
    flow = lambda t: W.spectral_flow_at(psih0, omega, t, n)
    o = O.Problem(g, flow, kappa=0.0, eta=0.0, kappah=1e-6, nkappah=4, dt=dt, stepper="ETDRK4")
    o.set_c(c0)
    o.stepforward(nsteps)
    p = TAD.Problem(dev, ptf.SpectralStreamfunctionFlow(psih0, omega), nx=n, κ=0.0, η=0.0,
                    κh=1e-6, nκh=4, dt=dt, stepper="ETDRK4")
    TAD.set_c_(p, c0)
    ptf.stepforward_(p, nsteps)
    assert rel(p.sol, o.sol) <= 1e-12
    p32 = TAD.Problem(dev, ptf.SpectralStreamfunctionFlow(psih0, omega), nx=n, κ=0.0, η=0.0,
                      κh=1e-6, nκh=4, dt=dt, stepper="ETDRK4", T=np.float32)
    TAD.set_c_(p32, c0)
    ptf.stepforward_(p32, nsteps)
    assert rel(p32.sol, o.sol) <= 1e-5


def test_host_time_varying_flow_parity(dev):
    """A non-uniform time-varying user flow, evaluated on the host at clock.t."""
    n, nsteps, dt = 64, 15, 0.01
    g = O.Grid(2, n)
    X, Y = W.grid_xyz(2, n)
    u = lambda x, y, t: np.cos(y) * (1 + 0.5 * np.sin(3 * t))
    v = lambda x, y, t: np.sin(x) * np.cos(2 * t)
    c0 = np.exp(-((X - 0.3) ** 2 + Y ** 2) / 0.5)
    o = O.Problem(g, lambda t: (u(X, Y, t), v(X, Y, t)), kappa=1e-3, dt=dt)
    o.set_c(c0)
    o.stepforward(nsteps)
    p = TAD.Problem(dev, ptf.TwoDAdvectingFlow(u=u, v=v, steadyflow=False), nx=n, κ=1e-3, dt=dt)
    TAD.set_c_(p, c0)
    ptf.stepforward_(p, nsteps)
    assert rel(p.sol, o.sol) <= 1e-12


def test_set_c_updatevars_roundtrip(dev):
    g = O.Grid(3, 32, ny=16, nz=64)
    rng = np.random.default_rng(3)
    c = rng.standard_normal(g.shape)
    p = TAD.Problem(dev, ptf.ThreeDAdvectingFlow(), nx=32, ny=16, nz=64)
    TAD.set_c_(p, c)
    assert rel(p.vars.c, c) <= 1e-14
    assert rel(p.sol, np.fft.rfftn(c, axes=(0, 1, 2))) <= 1e-14
    assert abs(TAD.mass(p) - c.sum()) <= 1e-12 * np.abs(c).sum()
    assert abs(TAD.variance(p) - c.var()) <= 1e-13 * c.var()
