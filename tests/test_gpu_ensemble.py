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
    """Graph replay (blocks of 16 steps with each size group in its own chain, then single
    steps) is bitwise the direct launches, across several calls with host work between."""
    dt = 5e-3
    members = _members()[:5]       # groups of 64^2, 128^2 and 256^2 members
    a = _ens(members, "fp64", dt)
    b = _ens(members, "fp64", dt)
    for n in (19, 17, 3, 32):      # 1 block + 3, 1 block + 1, 3 single steps, 2 blocks
        a.step(n, graph=True)
        b.step(n, graph=False)
        for i in range(len(members)):
            assert np.array_equal(a.get_sol(i), b.get_sol(i)), (n, i)
        assert a.clock(0) == b.clock(0)


def test_ensemble_member_equals_single_problem(dev, monkeypatch):
    """On the pass kernels (PTF_MEMBER_KERNEL=0) an ensemble member is bitwise the single
    problem; the whole-member kernel (the default for groups of small members) uses other
    radices, so it agrees to roundoff."""
    dt, nsteps = 5e-3, 10
    members = _members()[:4]
    monkeypatch.setenv("PTF_MEMBER_KERNEL", "0")
    e = _ens(members, "fp64", dt)
    e.step(nsteps)
    monkeypatch.delenv("PTF_MEMBER_KERNEL")
    em = _ens(members, "fp64", dt)
    assert any("member" in s for s in em.slots())
    em.step(nsteps)
    TAD = ptf.TracerAdvectionDiffusion
    for i, m in enumerate(members):
        u, v = m["flow"]
        p = TAD.Problem(dev, ptf.TwoDAdvectingFlow(u=lambda x, y: u, v=lambda x, y: v),
                        nx=m["n"], κ=m["kappa"], dt=dt)
        TAD.set_c_(p, m["c0"])
        ptf.stepforward_(p, nsteps)
        assert np.array_equal(p.sol, e.get_sol(i))
        assert rel(em.get_sol(i), p.sol) <= 1e-13, i


@pytest.mark.parametrize("stepper", ["RK4", "ETDRK4", "FilteredRK4"])
def test_ensemble_member_kernels_vs_oracle(dev, stepper):
    """Groups of 32^2, 64^2 and 128^2 members (whole-member step kernels: 32^2 and 64^2 in
    both forms, 128^2 in the mixed form) against the oracle: FP64 <= 1e-12, mixed <= 1e-5;
    and a time-varying flow (uploaded at each step's clock.t) on a member kernel."""
    dt, nsteps = 4e-3, 10
    members = []
    for i, n in enumerate([32, 64, 32, 128, 64, 128, 64]):
        members.append(W.ensemble_member(n, seed=2000 + i))
    kw = dict(order=6, innerK=0.5, outerK=1.0, tol=1e-12) if stepper == "FilteredRK4" else {}

    def moving(X, Y, t):
        return (0.4 * np.cos(t) - np.sin(X) * np.cos(Y), np.cos(X) * np.sin(Y) + 0.2 * np.sin(t))
    descs = [dict(n=m["n"], kappa=m["kappa"]) for m in members]
    descs[1]["flow"] = moving   # one 64^2 member with a time-varying flow
    for prec, bar in (("fp64", 1e-12), ("mixed", 1e-5)):
        print(f"member kernels {stepper} {prec}", flush=True)
        e = ptf.Ensemble(descs, dt=dt, precision=prec, stepper=stepper, **kw)
        names = e.slots()
        assert sum("member" in s for s in names) == (3 if prec == "mixed" else 2), names
        for i, m in enumerate(members):
            if i != 1:
                e.set_flow(i, *m["flow"])
            e.set_c(i, m["c0"])
        e.step(nsteps, graph=True)
        for i, m in enumerate(members):
            g = O.Grid(2, m["n"])
            if i == 1:
                X, Y = g.gridpoints()
                flow = (lambda t, X=X, Y=Y: moving(X, Y, t))
            else:
                flow = m["flow"]
            o = O.Problem(g, flow, kappa=m["kappa"], dt=dt, stepper=stepper, **kw)
            o.set_c(m["c0"])
            o.stepforward(nsteps)
            assert rel(e.get_sol(i), o.sol) <= bar, (prec, i, m["n"])


def test_ensemble_filtered_timevarying_clock_and_diagnostics(dev):
    """Members with FilteredRK4 (custom filter), a time-varying flow evaluated at each
    member's clock.t, and per-member clocks and FP64 diagnostics, against the oracle."""
    dt, nsteps = 5e-3, 12
    base = _members()[:4]

    def moving(U):
        # a uniformly translating cellular flow: u(x, y, t) = U cos(t) + cellular
        return lambda X, Y, t: (U * np.cos(t) - np.sin(X) * np.cos(Y), np.cos(X) * np.sin(Y))
    members = []
    for k, m in enumerate(base):
        mm = dict(n=m["n"], kappa=m["kappa"])
        if k % 2:
            mm["flow"] = moving(0.3 + 0.1 * k)
        members.append(mm)
    filt = dict(order=6, innerK=0.5, outerK=1.0, tol=1e-12)
    e = ptf.Ensemble(members, dt=dt, precision="fp64", stepper="FilteredRK4", **filt)
    for i, m in enumerate(base):
        if i % 2 == 0:
            e.set_flow(i, *m["flow"])
        e.set_c(i, m["c0"])
    e.step(nsteps, graph=True)
    for i, m in enumerate(base):
        g = O.Grid(2, m["n"])
        if i % 2:
            f = members[i]["flow"]
            X, Y = g.gridpoints()
            flow = (lambda t, f=f, X=X, Y=Y: f(X, Y, t))
        else:
            flow = m["flow"]
        o = O.Problem(g, flow, kappa=m["kappa"], dt=dt, stepper="FilteredRK4", **filt)
        o.set_c(m["c0"])
        o.stepforward(nsteps)
        assert rel(e.get_sol(i), o.sol) <= 1e-12, i
        t, s = e.clock(i)
        assert s == nsteps and t == o.clock.t
        mass, var = e.diagnostics(i)
        assert abs(mass - O.mass(g, o.sol)) <= 1e-12 * np.abs(o.sol).max()
        assert abs(var - O.variance(g, o.sol)) <= 1e-12 * O.variance(g, o.sol)
    e.set_clock(0.0, 0)
    assert e.clock(0) == (0.0, 0)
