"""Stored fixtures (tests/golden/, written by make_golden.py from the oracle)."""
import os

import numpy as np
import pytest

from golden.make_golden import CASES, build

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


@pytest.mark.parametrize("name", sorted(CASES))
def test_golden_oracle_reproduces(name):
    ref = load(name)
    d = build(name, CASES[name])
    for k in ("flows", "c0", "L"):
        assert np.array_equal(d[k], ref[k]), k
    assert np.linalg.norm(d["sol"] - ref["sol"]) <= 1e-14 * np.linalg.norm(ref["sol"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_golden_device(dev, name):
    import passivetracerflows as ptf
    TAD = ptf.TracerAdvectionDiffusion
    c, ref = CASES[name], load(name)
    nd = c["ndim"]
    nx, ny, nz = c["n"]
    fl = [lambda *a, i=i: ref["flows"][i] for i in range(nd)]
    if nd == 1:
        gr = ptf.OneDGrid(dev, nx=nx)
        pr = TAD.ConstDiffSteadyFlowParams(c["kappa"], c["kappah"], c["nkappah"], fl[0], gr)
    elif nd == 2:
        gr = ptf.TwoDGrid(dev, nx=nx, ny=ny)
        pr = TAD.ConstDiffSteadyFlowParams(c["kappa"], c["eta"], c["kappah"], c["nkappah"],
                                           fl[0], fl[1], gr)
    else:
        gr = ptf.ThreeDGrid(dev, nx=nx, ny=ny, nz=nz)
        pr = TAD.ConstDiffSteadyFlowParams(c["kappa"], c["eta"], c["iota"], c["kappah"],
                                           c["nkappah"], fl[0], fl[1], fl[2], gr)
    prob = ptf.FourierFlows.Problem(TAD.Equation(pr, gr), c["stepper"], c["dt"], gr,
                                    TAD.Vars(gr), pr)
    assert np.allclose(prob.eqn.L, ref["L"], rtol=1e-15, atol=0)
    TAD.set_c_(prob, ref["c0"])
    ptf.stepforward_(prob, c["nsteps"])
    assert np.linalg.norm(prob.sol - ref["sol"]) <= 1e-12 * np.linalg.norm(ref["sol"])
