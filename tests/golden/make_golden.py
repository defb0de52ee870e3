"""Write the small golden fixtures of tests/golden/ from the FP64 oracle.

    python tests/golden/make_golden.py

Each fixture holds the inputs (grid sizes, parameters, flows, c0) and the
expected spectral state after a few steps.  The reference (Julia) cannot run in
this image, so these are oracle outputs: they guard the oracle against drift
(test_golden_oracle_reproduces) and give the GPU tests stored expectations.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import tad as O  # noqa: E402

CASES = {
    "rk4_2d_32x16": dict(ndim=2, n=(32, 16, 1), kappa=0.01, eta=0.02, kappah=0.0, nkappah=0,
                         dt=0.02, nsteps=6, stepper="RK4", seed=11),
    "etdrk4_3d_16": dict(ndim=3, n=(16, 16, 16), kappa=0.03, eta=0.02, iota=0.01, kappah=1e-4,
                         nkappah=2, dt=0.05, nsteps=4, stepper="ETDRK4", seed=12),
    "rk4_1d_64_hyper": dict(ndim=1, n=(64, 1, 1), kappa=0.0, kappah=1e-5, nkappah=2, dt=0.01,
                            nsteps=8, stepper="RK4", seed=13),
}


def build(name, c):
    rng = np.random.default_rng(c["seed"])
    nd = c["ndim"]
    nx, ny, nz = c["n"]
    g = O.Grid(nd, nx, ny=ny if nd >= 2 else None, nz=nz if nd == 3 else None)
    flows = tuple(np.cos(rng.integers(1, 3) * np.stack(g.gridpoints())[i % nd] + rng.uniform(0, 6))
                  * rng.uniform(0.2, 0.8) for i in range(nd))
    c0 = rng.standard_normal(g.shape)
    c0 = O.irfft(g, O.rfft(g, c0) * np.exp(-g.Krsq / 10.0))
    p = O.Problem(g, flows, kappa=c["kappa"], eta=c.get("eta"), iota=c.get("iota"),
                  kappah=c["kappah"], nkappah=c["nkappah"], dt=c["dt"], stepper=c["stepper"])
    p.set_c(c0)
    p.stepforward(c["nsteps"])
    return dict(flows=np.stack(flows), c0=c0, sol=p.sol, L=p.L)


if __name__ == "__main__":
    for name, c in CASES.items():
        d = build(name, c)
        np.savez(os.path.join(HERE, name + ".npz"), **d)
        print("wrote", name)
