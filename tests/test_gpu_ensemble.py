"""Ensemble (batched) form: members of several grid sizes advanced together.

Bars: FP64 members match their own oracle runs to relative L2 <= 1e-12; the
FP32 (mixed) members match the FP64 oracle to <= 1e-5; hipGraph replay is
bitwise identical to direct launches.
"""
import numpy as np
import pytest

import passivetracerflows as ptf
from oracle import tad as O
from passivetracerflows import workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def _members():
    out = []
    for i, n in enumerate([64, 64, 128, 64, 256, 128, 64, 512]):
        m = W.ensemble_member(n, seed=1000 + i)
        out.append(m)
    return out


def _ens(members, precision, dt):
    e = ptf.Ensemble([dict(n=m["n"], kappa=m["kappa"]) for m in members], dt=dt,
                     precision=precision)
    for i, m in enumerate(members):
        e.set_flow(i, *m["flow"])
        e.set_c(i, m["c0"])
    return e


def _oracle(m, dt, nsteps):
    g = O.Grid(2, m["n"])
    p = O.Problem(g, m["flow"], kappa=m["kappa"], dt=dt)
    p.set_c(m["c0"])
    p.stepforward(nsteps)
    return p.sol


def test_ensemble_fp64_and_mixed_vs_oracle(dev):
    dt, nsteps = 5e-3, 20
    members = _members()
    print(f"ensemble of {len(members)}: sizes {[m['n'] for m in members]}", flush=True)
    e64 = _ens(members, "fp64", dt)
    e64.step(nsteps, graph=True)
    e32 = _ens(members, "mixed", dt)
    e32.step(nsteps, graph=True)
    for i, m in enumerate(members):
        ref = _oracle(m, dt, nsteps)
        assert rel(e64.get_sol(i), ref) <= 1e-12, i
        assert rel(e32.get_sol(i), ref) <= 1e-5, i
        c = e64.get_c(i)
        assert rel(c, O.irfft(O.Grid(2, m["n"]), ref)) <= 1e-12


def test_ensemble_graph_equals_direct(dev):
    dt = 5e-3
    members = _members()[:5]
    a = _ens(members, "fp64", dt)
    b = _ens(members, "fp64", dt)
    a.step(7, graph=True)
    b.step(7, graph=False)
    for i in range(len(members)):
        assert np.array_equal(a.get_sol(i), b.get_sol(i))
    # the graph replays across several calls, with host work in between
    a.step(3, graph=True)
    b.step(3, graph=False)
    for i in range(len(members)):
        assert np.array_equal(a.get_sol(i), b.get_sol(i))


def test_ensemble_member_equals_single_problem(dev):
    dt, nsteps = 5e-3, 10
    members = _members()[:4]
    e = _ens(members, "fp64", dt)
    e.step(nsteps)
    TAD = ptf.TracerAdvectionDiffusion
    for i, m in enumerate(members):
        u, v = m["flow"]
        p = TAD.Problem(dev, ptf.TwoDAdvectingFlow(u=lambda x, y: u, v=lambda x, y: v),
                        nx=m["n"], κ=m["kappa"], dt=dt)
        TAD.set_c_(p, m["c0"])
        ptf.stepforward_(p, nsteps)
        assert np.array_equal(p.sol, e.get_sol(i))
