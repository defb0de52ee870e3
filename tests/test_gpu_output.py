"""GPU-resident diagnostics and snapshots (ptf_step_record) against the per-step
path and the oracle, and the reference's 1D example output loop
(examples/onedim_gaussiandiffusion.jl:44-95) end to end.

Bars: device records equal the per-step ptf_diagnostics/updatevars values of an
identical problem (same kernels: exact); fields against the FP64 oracle <= 1e-12.
"""
import numpy as np
import pytest

import passivetracerflows as ptf
from oracle import tad as O
from passivetracerflows import workloads as W

pytestmark = pytest.mark.gpu
TAD = ptf.TracerAdvectionDiffusion


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def _cellular(dev, n, T=np.float64):
    cfg = W.cellular_2d(n)
    u, v = cfg["flow"]
    p = TAD.Problem(dev, ptf.TwoDAdvectingFlow(u=lambda x, y: u, v=lambda x, y: v), nx=n,
                    κ=cfg["kappa"], dt=cfg["dt"], T=T)
    TAD.set_c_(p, cfg["c0"])
    return cfg, p


@pytest.mark.parametrize("n", [64, 128])
def test_record_matches_per_step_path(dev, n):
    print(f"record n={n}", flush=True)
    cfg, a = _cellular(dev, n)
    _, b = _cellular(dev, n)
    rec = ptf.record_(a, 23, 5)
    assert list(rec["step"]) == [5, 10, 15, 20]
    assert np.allclose(rec["t"], 0.01 * rec["step"], rtol=0, atol=1e-12)
    assert a.clock.step == 23
    for i in range(4):
        ptf.stepforward_(b, 5)
        assert TAD.mass(b) == rec["mass"][i]
        assert TAD.variance(b) == rec["variance"][i]
        TAD.updatevars_(b)
        assert np.array_equal(b.vars.c, rec["c"][i])
    # the last snapshot against the oracle
    g = O.Grid(2, n)
    o = O.Problem(g, cfg["flow"], kappa=cfg["kappa"], dt=cfg["dt"])
    o.set_c(cfg["c0"])
    o.stepforward(20)
    assert rel(rec["c"][3], o.updatevars()) <= 1e-12


def test_stepforward_with_diagnostics(dev):
    cfg, p = _cellular(dev, 64)
    dm = ptf.Diagnostic(TAD.mass, p, freq=4, nsteps=12)
    dv = ptf.Diagnostic(TAD.variance, p, freq=4, nsteps=12)
    ptf.stepforward_(p, [dm, dv], 12)
    assert dm.steps == [0, 4, 8, 12] and dv.steps == [0, 4, 8, 12]
    # a host diagnostic takes the per-step path and agrees with the device records
    _, q = _cellular(dev, 64)
    dh = ptf.Diagnostic(lambda pr: TAD.variance(pr), q, freq=4, nsteps=12)
    ptf.stepforward_(q, dh, 12)
    assert dh.data == dv.data
    g = O.Grid(2, 64)
    o = O.Problem(g, cfg["flow"], kappa=cfg["kappa"], dt=cfg["dt"])
    o.set_c(cfg["c0"])
    o.stepforward(12)
    assert abs(dv.data[-1] - O.variance(g, o.sol)) <= 1e-12 * O.variance(g, o.sol)
    assert abs(dm.data[-1] - O.mass(g, o.sol)) <= 1e-12 * np.abs(o.sol).max()


def test_reference_1d_example_output_loop(dev, tmp_path):
    """examples/onedim_gaussiandiffusion.jl: n=128, dt=0.02, kappa=0.01, u=0.05, Gaussian
    spread 0.15; save every 50 steps (here 400 of its 5000 steps)."""
    n, dt, kappa, nsteps, every = 128, 0.02, 0.01, 400, 50
    prob = TAD.Problem(dev, ptf.OneDAdvectingFlow(u=lambda x: 0.05 + 0 * x), nx=n, Lx=2 * np.pi,
                       κ=kappa, dt=dt, stepper="RK4")
    x = prob.grid.x
    c0 = np.exp(-x ** 2 / (2 * 0.15 ** 2))
    TAD.set_c_(prob, c0)

    def get_concentration(pr):
        TAD.updatevars_(pr)
        return pr.vars.c

    out = ptf.Output(prob, str(tmp_path / "advection-diffusion1D.jld2"),
                     ("concentration", get_concentration))
    ptf.saveproblem(out)
    while prob.clock.step <= nsteps:
        if prob.clock.step % every == 0:
            ptf.saveoutput(out)
        ptf.stepforward_(prob)
    f = ptf.jldopen(out.path)
    iters = sorted(int(k) for k in f.keys("snapshots/t"))
    assert iters == list(range(0, nsteps + 1, every))
    assert f["grid/Lx"] == 2 * np.pi and np.array_equal(f["grid/x"], x)
    g = O.Grid(1, n)
    o = O.Problem(g, (0.05 + 0 * x,), kappa=kappa, dt=dt)
    o.set_c(c0)
    for i in iters[1:]:
        o.stepforward(every)
        assert abs(f[f"snapshots/t/{i}"] - i * dt) < 1e-12
        assert rel(f[f"snapshots/concentration/{i}"], o.updatevars()) <= 1e-12
    # the GPU-resident path writes the same snapshots
    p2 = TAD.Problem(dev, ptf.OneDAdvectingFlow(u=lambda x: 0.05 + 0 * x), nx=n, κ=kappa, dt=dt)
    TAD.set_c_(p2, c0)
    out2 = ptf.Output(p2, str(tmp_path / "recorded.jld2"))
    ptf.save_records(out2, ptf.record_(p2, nsteps, every))
    f2 = ptf.jldopen(out2.path)
    for i in iters[1:]:
        assert np.array_equal(f2[f"snapshots/concentration/{i}"], f[f"snapshots/concentration/{i}"])
