"""The two large workloads at their full step counts against the FP64 oracle.

The oracle needs tens of minutes for 3D 256^3 x 500 RK4 steps and for 2D 2048^2 x
1000 ETDRK4 steps, so its final states are committed as fingerprints
(tests/golden/make_fullrun.py, tests/fingerprint.py): the |k| <= 16 modes, the norm,
mass, variance and four seeded projections over the whole spectrum.  The device runs
the same workloads and every fingerprint entry must agree: FP64 within 1e-12 (the
relative L2 bar of the plain form), the mixed form within 1e-5.  The measured errors
are printed.
"""
import os

import numpy as np
import pytest

import passivetracerflows as ptf
from passivetracerflows import workloads as W
import fingerprint as F

pytestmark = pytest.mark.gpu
TAD = ptf.TracerAdvectionDiffusion
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with np.load(os.path.join(GOLDEN, f"fullrun_{name}.npz")) as z:
        return {k: z[k] for k in z.files}


def _fl(fields):
    return [lambda *a, f=f: f for f in fields]


def _check(p, ref, bar, label):
    npts = int(ref["npoints"])
    err = F.compare(p.sol, ref, npts)
    print(f"{label}: " + ", ".join(f"{k} {v:.2e}" for k, v in err.items()), flush=True)
    assert p.clock.step == int(ref["nsteps"])
    assert abs(p.clock.t - float(ref["t"])) <= 1e-12 * float(ref["t"])
    for k, v in err.items():
        assert v <= bar, (label, k, v)
    return err


@pytest.mark.parametrize("T,bar,tune", [(np.float64, 1e-12, False), (np.float64, 1e-12, True),
                                        (np.float32, 1e-5, True)])
def test_abc3d_256_full_length(dev, T, bar, tune):
    """3D ABC flow 256^3, kappa 1e-3, RK4 dt 5e-3, 500 steps."""
    ref = _load("abc3d")
    cfg = W.abc_3d(256)
    print(f"abc3d 256^3 x {cfg['nsteps']} {T.__name__} tune={tune}", flush=True)
    p = TAD.Problem(dev, ptf.ThreeDAdvectingFlow(*_fl(cfg["flow"])), nx=256, κ=cfg["kappa"],
                    dt=cfg["dt"], T=T)
    TAD.set_c_(p, cfg["c0"])
    # the initial state is the oracle's too
    e0 = F.compare(p.sol, {k[5:]: v for k, v in ref.items() if k.startswith("init_")},
                   int(ref["npoints"]))
    assert max(e0.values()) <= (1e-14 if T is np.float64 else 1e-6), e0
    if tune:
        p.autotune()
    ptf.stepforward_(p, cfg["nsteps"])
    _check(p, ref, bar, f"abc3d {T.__name__} tune={tune}")


@pytest.mark.parametrize("T,bar,tune", [(np.float64, 1e-12, False), (np.float64, 1e-12, True),
                                        (np.float32, 1e-5, True)])
def test_turbulent2d_2048_full_length(dev, T, bar, tune):
    """2D random-phase flow 2048^2 at clock.t, hyperdiffusion 1e-12 (order 4), ETDRK4 dt 2e-3,
    1000 steps."""
    ref = _load("turb2d")
    cfg = W.turbulent_2d(2048)
    print(f"turb2d 2048^2 x {cfg['nsteps']} {T.__name__} tune={tune}", flush=True)
    p = TAD.Problem(dev, ptf.SpectralStreamfunctionFlow(cfg["psih0"], cfg["omega"]), nx=2048,
                    κ=0.0, η=0.0, κh=cfg["kappah"], nκh=cfg["nkappah"], dt=cfg["dt"],
                    stepper="ETDRK4", T=T)
    TAD.set_c_(p, cfg["c0"])
    if tune:
        p.autotune()
    ptf.stepforward_(p, cfg["nsteps"])
    _check(p, ref, bar, f"turb2d {T.__name__} tune={tune}")
