"""Fingerprints of the FP64 oracle after the full step counts of the two large workloads.

    python tests/golden/make_fullrun.py abc3d      # 3D ABC flow 256^3, RK4, 500 steps
    python tests/golden/make_fullrun.py turb2d     # 2D random-phase flow 2048^2, ETDRK4, 1000 steps

The oracle needs tens of minutes for each, far too long for a test, so the final
state is stored as a fingerprint (tests/fingerprint.py): the |k| <= 16 modes, the
norm, mass, variance and four seeded projections over the whole spectrum.  The
GPU tests run the same workloads at full length and compare against these.
Written to tests/golden/fullrun_<name>.npz.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "passivetracerflows.jl_amd"))
sys.path.insert(0, os.path.dirname(HERE))
from oracle import tad as O  # noqa: E402
from passivetracerflows import workloads as W  # noqa: E402
import fingerprint as F  # noqa: E402


def _run(o, nsteps, label, every=25):
    t0 = time.time()
    while o.clock.step < nsteps:
        o.stepforward(min(every, nsteps - o.clock.step))
        el = time.time() - t0
        print(f"{label}: step {o.clock.step}/{nsteps}  {el:.0f} s", flush=True)


def abc3d():
    cfg = W.abc_3d(256)
    g = O.Grid(3, 256)
    o = O.Problem(g, cfg["flow"], kappa=cfg["kappa"], dt=cfg["dt"])
    o.set_c(cfg["c0"])
    fp0 = F.fingerprint(o.sol, g.npoints)
    _run(o, cfg["nsteps"], "abc3d")
    return g, o, fp0, cfg


def turb2d():
    cfg = W.turbulent_2d(2048)
    n = 2048
    g = O.Grid(2, n)
    flow = lambda t: W.spectral_flow_at(cfg["psih0"], cfg["omega"], t, n)
    o = O.Problem(g, flow, kappa=0.0, eta=0.0, kappah=cfg["kappah"], nkappah=cfg["nkappah"],
                  dt=cfg["dt"], stepper="ETDRK4")
    o.set_c(cfg["c0"])
    fp0 = F.fingerprint(o.sol, g.npoints)
    _run(o, cfg["nsteps"], "turb2d")
    return g, o, fp0, cfg


if __name__ == "__main__":
    name = sys.argv[1]
    g, o, fp0, cfg = {"abc3d": abc3d, "turb2d": turb2d}[name]()
    fp = F.fingerprint(o.sol, g.npoints)
    out = os.path.join(HERE, f"fullrun_{name}.npz")
    np.savez(out, nsteps=cfg["nsteps"], t=o.clock.t, npoints=g.npoints,
             **{k: v for k, v in fp.items()}, **{"init_" + k: v for k, v in fp0.items()})
    print("wrote", out, "norm", fp["norm"], "mass", fp["mass"], "var", fp["var"], flush=True)
