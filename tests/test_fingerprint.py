"""The long-run fingerprints (tests/fingerprint.py) detect the errors they are meant to.

CPU tests: on small oracle runs a planted mistake (in the oracle, never in the
product) moves some fingerprint entry far beyond the 1e-12 bar, roundoff-level
noise does not, and the projection entry tracks the relative L2 error of the
whole field.  The committed full-length fixtures start from the workload
generators' own initial states.
"""
import os

import numpy as np
import pytest

from oracle import tad as O
from passivetracerflows import workloads as W
import fingerprint as F

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _run(nsteps=5, flow_scale=(1.0, 1.0, 1.0), kappa=0.02, sol_hook=None):
    g = O.Grid(3, 16)
    X, Y, Z = g.gridpoints()
    flow = (flow_scale[0] * (np.sin(Z) + np.cos(Y)), flow_scale[1] * (np.sin(X) + np.cos(Z)),
            flow_scale[2] * (np.sin(Y) + np.cos(X)))
    o = O.Problem(g, flow, kappa=kappa, dt=0.01)
    rng = np.random.default_rng(1)
    c0 = O.irfft(g, O.rfft(g, rng.standard_normal(g.shape)) * np.exp(-g.Krsq / 6))
    o.set_c(c0)
    o.stepforward(nsteps)
    return g, (sol_hook(o.sol) if sol_hook else o.sol)


def test_planted_mistakes_are_flagged():
    g, ref = _run()
    fp = F.fingerprint(ref, g.npoints)
    assert max(F.compare(ref, fp, g.npoints).values()) == 0.0
    # a flow component 1e-8 too strong; a different diffusivity; one step short
    for bad in (_run(flow_scale=(1.0, 1.0, 1.0 + 1e-8))[1], _run(kappa=0.0200001)[1],
                _run(nsteps=4)[1]):
        err = F.compare(bad, fp, g.npoints)
        assert max(err.values()) > 1e-10, err
    # a single mode of the whole spectrum wrong by 1e-9 of the norm
    bad = ref.copy()
    bad[7, 9, 5] += 1e-9 * np.linalg.norm(ref)
    assert F.compare(bad, fp, g.npoints)["proj"] > 1e-11


def test_roundoff_passes_and_projection_tracks_rel_l2():
    g, ref = _run()
    fp = F.fingerprint(ref, g.npoints)
    rng = np.random.default_rng(3)
    for eps in (1e-15, 1e-9):
        noise = rng.standard_normal(ref.shape) + 1j * rng.standard_normal(ref.shape)
        pert = ref + eps * np.linalg.norm(ref) * noise / np.linalg.norm(noise)
        err = F.compare(pert, fp, g.npoints)
        assert err["proj"] <= 5 * eps and err["proj"] >= eps / 50, err
        assert err["low"] <= 50 * eps and err["norm"] <= 5 * eps, err
        if eps < 1e-12:
            assert max(err.values()) <= 1e-13, err


@pytest.mark.parametrize("name", ["abc3d", "turb2d"])
def test_fullrun_fixture_initial_state(name):
    """The stored initial fingerprint is the oracle's transform of the workload's c0."""
    with np.load(os.path.join(GOLDEN, f"fullrun_{name}.npz")) as z:
        ref = {k[5:]: z[k] for k in z.files if k.startswith("init_")}
        npts = int(z["npoints"])
        assert int(z["nsteps"]) == (500 if name == "abc3d" else 1000)
    if name == "abc3d":
        c0, g = W.abc_3d(256)["c0"], O.Grid(3, 256)
    else:
        c0, g = W.turbulent_2d(2048)["c0"], O.Grid(2, 2048)
    assert g.npoints == npts
    err = F.compare(O.rfft(g, c0), ref, npts)
    assert max(err.values()) <= 1e-14, err
