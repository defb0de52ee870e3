"""The reference's on-path analytic-solution tests, restated as data.

PassiveTracerFlows.jl test/runtests.jl:26-56 runs these with the tolerances
of test/test_traceradvectiondiffusion.jl (T):

  constvel1D/2D/3D          T:8-33, T:79-106, T:155-182     c0(x - U t)
  timedependentvel1D/2D/3D  T:42-70, T:116-145, T:192-224   c0(x - U t^2/2)  (flow at clock.t)
  diffusion1D/2D/3D         T:232-264, T:272-301, T:363-392 heat kernel (steady and unsteady)
  hyperdiffusion            T:400-439                       heat kernel, kappah = 0.01, nkappah = 1

Julia's isapprox(a, b; rtol) is  norm(a - b) <= rtol * max(norm(a), norm(b)).
(test_diffusion_multilayerqg, T:303-355, needs GeophysicalFlows and is not on this path.)
"""
import numpy as np

TWO_PI = 2 * np.pi


def isapprox(a, b, rtol):
    return float(np.linalg.norm(np.ravel(a - b))) <= rtol * max(np.linalg.norm(np.ravel(a)),
                                                                np.linalg.norm(np.ravel(b)))


def _pts(n, ndim):
    start = np.longdouble(-np.pi)
    step = np.longdouble(TWO_PI / n)
    x = (start + np.arange(n, dtype=np.longdouble) * step).astype(np.float64)
    if ndim == 1:
        return (x,)
    if ndim == 2:
        Y, X = np.meshgrid(x, x, indexing="ij")
        return X, Y
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    return X, Y, Z


def case(name):
    """Returns dict: ndim, nx, kappa, kappah, nkappah, dt, nsteps, steady, flow (functions of
    (pts..., t) or of pts...), c0, cfinal, rtol, lowlevel."""
    stepper = "RK4"
    if name == "constvel1D":
        nx, dt, nsteps, U = 128, 1e-2, 40, 0.05
        (x,) = _pts(nx, 1)
        f = lambda x: 0.1 * np.exp(-x ** 2 / (2 * 0.1 ** 2))
        t = nsteps * dt
        return dict(ndim=1, nx=nx, kappa=0.0, dt=dt, nsteps=nsteps, steady=True,
                    flow=(lambda x: U + 0 * x,), c0=f(x), cfinal=f(x - U * t),
                    rtol=nx * nsteps * 1e-12, stepper=stepper)
    if name == "timedependentvel1D":
        nx, dt, tfinal, U = 128, 0.002, 0.1, 0.05
        nsteps = round(tfinal / dt)
        (x,) = _pts(nx, 1)
        f = lambda x: 0.1 * np.exp(-x ** 2 / (2 * 0.2 ** 2))
        t = nsteps * dt
        return dict(ndim=1, nx=nx, kappa=0.0, dt=dt, nsteps=nsteps, steady=False,
                    flow=(lambda x, t: U * t + U * dt / 2 + 0 * x,), c0=f(x),
                    cfinal=f(x - 0.5 * U * t ** 2), rtol=nx * nsteps * 1e-12, stepper=stepper)
    if name == "constvel2D":
        nx, dt, nsteps, U, V = 128, 1e-2, 40, 0.2, 0.1
        X, Y = _pts(nx, 2)
        f = lambda x, y: 0.1 * np.exp(-(x ** 2 + y ** 2) / (2 * 0.1 ** 2))
        t = nsteps * dt
        return dict(ndim=2, nx=nx, kappa=0.0, dt=dt, nsteps=nsteps, steady=True,
                    flow=(lambda x, y: U + 0 * x, lambda x, y: V + 0 * x), c0=f(X, Y),
                    cfinal=f(X - U * t, Y - V * t), rtol=nx * nx * nsteps * 1e-12, stepper=stepper)
    if name == "timedependentvel2D":
        nx, dt, tfinal, U, av = 128, 0.002, 0.1, 0.5, 0.5
        nsteps = round(tfinal / dt)
        X, Y = _pts(nx, 2)
        f = lambda x, y: 0.1 * np.exp(-(x ** 2 + y ** 2) / (2 * 0.2 ** 2))
        t = nsteps * dt
        return dict(ndim=2, nx=nx, kappa=0.0, dt=dt, nsteps=nsteps, steady=False,
                    flow=(lambda x, y, t: U + 0 * x, lambda x, y, t: av * t + av * dt / 2 + 0 * x),
                    c0=f(X, Y), cfinal=f(X - U * t, Y - 0.5 * av * t ** 2),
                    rtol=nx * nx * nsteps * 1e-12, stepper=stepper)
    if name == "constvel3D":
        nx, dt, nsteps, U, V, W = 128, 1e-2, 40, 0.2, 0.1, 0.05
        X, Y, Z = _pts(nx, 3)
        f = lambda x, y, z: 0.1 * np.exp(-(x ** 2 + y ** 2 + z ** 2) / (2 * 0.1 ** 2))
        t = nsteps * dt
        return dict(ndim=3, nx=nx, kappa=0.0, dt=dt, nsteps=nsteps, steady=True,
                    flow=(lambda x, y, z: U + 0 * x, lambda x, y, z: V + 0 * x,
                          lambda x, y, z: W + 0 * x),
                    c0=f(X, Y, Z), cfinal=f(X - U * t, Y - V * t, Z - W * t),
                    rtol=nx ** 3 * nsteps * 1e-12, stepper=stepper)
    if name == "timedependentvel3D":
        nx, dt, tfinal, U, av, W = 128, 0.002, 0.1, 0.5, 0.5, 0.5
        nsteps = round(tfinal / dt)
        X, Y, Z = _pts(nx, 3)
        f = lambda x, y, z: 0.1 * np.exp(-(x ** 2 + y ** 2 + z ** 2) / (2 * 0.2 ** 2))
        t = nsteps * dt
        return dict(ndim=3, nx=nx, kappa=0.0, dt=dt, nsteps=nsteps, steady=False,
                    flow=(lambda x, y, z, t: U + 0 * x,
                          lambda x, y, z, t: av * t + av * dt / 2 + 0 * x,
                          lambda x, y, z, t: W + 0 * x),
                    c0=f(X, Y, Z), cfinal=f(X - U * t, Y - 0.5 * av * t ** 2, Z - W * t),
                    rtol=nx ** 3 * nsteps * 1e-12, stepper=stepper)
    if name.startswith("diffusion1D"):
        steady = name.endswith("steady")
        nx, kappa, dt, tfinal = 128, 0.01, 0.005, 0.1
        nsteps = round(tfinal / dt)
        (x,) = _pts(nx, 1)
        s0 = 0.1
        sig = lambda t: np.sqrt(2 * kappa * t + s0)
        f = lambda x, t: (0.1 / sig(t)) * np.exp(-x ** 2 / (2 * sig(t) ** 2))
        t = nsteps * dt
        zero = (lambda x: 0 * x) if steady else (lambda x, t: 0 * x)
        return dict(ndim=1, nx=nx, kappa=kappa, dt=dt, nsteps=nsteps, steady=steady, flow=(zero,),
                    c0=f(x, 0.0), cfinal=f(x, t), rtol=nx * nsteps * 1e-12, stepper=stepper)
    if name.startswith("diffusion2D"):
        steady = name.endswith("steady")
        nx, kappa, dt, tfinal, s = 128, 0.01, 0.005, 0.1, 0.1
        nsteps = round(tfinal / dt)
        X, Y = _pts(nx, 2)
        t = nsteps * dt
        st = np.sqrt(2 * kappa * t + s ** 2)
        zero = (lambda x, y: 0 * x) if steady else (lambda x, y, t: 0 * x)
        return dict(ndim=2, nx=nx, kappa=kappa, dt=dt, nsteps=nsteps, steady=steady,
                    flow=(zero, zero), c0=0.1 * np.exp(-(X ** 2 + Y ** 2) / (2 * s ** 2)),
                    cfinal=0.1 * (s ** 2 / st ** 2) * np.exp(-(X ** 2 + Y ** 2) / (2 * st ** 2)),
                    rtol=nx * nx * nsteps * 1e-12, stepper=stepper)
    if name.startswith("diffusion3D"):
        steady = name.endswith("steady")
        nx, kappa, dt, tfinal, s = 64, 0.01, 0.005, 0.1, 0.1
        nsteps = round(tfinal / dt)
        X, Y, Z = _pts(nx, 3)
        t = nsteps * dt
        st = np.sqrt(2 * kappa * t + s ** 2)
        R2 = X ** 2 + Y ** 2 + Z ** 2
        zero = (lambda x, y, z: 0 * x) if steady else (lambda x, y, z, t: 0 * x)
        return dict(ndim=3, nx=nx, kappa=kappa, dt=dt, nsteps=nsteps, steady=steady,
                    flow=(zero, zero, zero), c0=0.1 * np.exp(-R2 / (2 * s ** 2)),
                    cfinal=0.1 * (s / st) ** 3 * np.exp(-R2 / (2 * st ** 2)),
                    rtol=nx ** 3 * nsteps * 1e-10, stepper=stepper)
    if name == "hyperdiffusion":
        nx, kh, dt, tfinal, s = 128, 0.01, 0.005, 0.1, 0.1
        nsteps = round(tfinal / dt)
        X, Y = _pts(nx, 2)
        t = nsteps * dt
        st = np.sqrt(2 * kh * t + s ** 2)
        return dict(ndim=2, nx=nx, kappa=0.0, kappah=kh, nkappah=1, dt=dt, nsteps=nsteps,
                    steady=True, lowlevel=True, flow=(lambda x, y: 0 * x, lambda x, y: 0 * x),
                    c0=0.1 * np.exp(-(X ** 2 + Y ** 2) / (2 * s ** 2)),
                    cfinal=0.1 * s ** 2 / st ** 2 * np.exp(-(X ** 2 + Y ** 2) / (2 * st ** 2)),
                    rtol=nx * nx * nsteps * 1e-12, stepper=stepper)
    raise KeyError(name)


NAMES = ["constvel1D", "timedependentvel1D", "constvel2D", "timedependentvel2D", "constvel3D",
         "timedependentvel3D", "diffusion1D_steady", "diffusion1D_unsteady", "diffusion2D_steady",
         "diffusion2D_unsteady", "diffusion3D_steady", "diffusion3D_unsteady", "hyperdiffusion"]


def run_oracle(c):
    """Run a case through the oracle restatement; returns final c."""
    from oracle import tad as O
    g = O.Grid(c["ndim"], c["nx"])
    pts = _pts(c["nx"], c["ndim"])
    if c["steady"]:
        flow = tuple(np.broadcast_to(f(*pts), g.shape) for f in c["flow"])
    else:
        flow = lambda t: tuple(np.broadcast_to(f(*pts, t), g.shape) for f in c["flow"])
    p = O.Problem(g, flow, kappa=c["kappa"], kappah=c.get("kappah", 0.0),
                  nkappah=c.get("nkappah", 0), dt=c["dt"], stepper=c["stepper"])
    p.set_c(c["c0"])
    p.stepforward(c["nsteps"])
    return p.updatevars()


def run_device(dev, c, T=np.float64):
    """Run a case through the device API with the reference's call sequence."""
    import passivetracerflows as ptf
    TAD = ptf.TracerAdvectionDiffusion
    nd, nx = c["ndim"], c["nx"]
    if c.get("lowlevel"):
        # T:413-423: params -> Equation -> FourierFlows.Problem
        gr = ptf.TwoDGrid(dev, nx=nx, Lx=TWO_PI, T=T)
        vs = TAD.Vars(gr)
        pr = TAD.ConstDiffSteadyFlowParams(c["kappa"], c["kappa"], c["kappah"], c["nkappah"],
                                           c["flow"][0], c["flow"][1], gr)
        eq = TAD.Equation(pr, gr)
        prob = ptf.FourierFlows.Problem(eq, c["stepper"], c["dt"], gr, vs, pr)
    else:
        if nd == 1:
            fl = ptf.OneDAdvectingFlow(u=c["flow"][0], steadyflow=c["steady"])
        elif nd == 2:
            fl = ptf.TwoDAdvectingFlow(u=c["flow"][0], v=c["flow"][1], steadyflow=c["steady"])
        else:
            fl = ptf.ThreeDAdvectingFlow(u=c["flow"][0], v=c["flow"][1], w=c["flow"][2],
                                         steadyflow=c["steady"])
        prob = TAD.Problem(dev, fl, nx=nx, Lx=TWO_PI, κ=c["kappa"], dt=c["dt"],
                           stepper=c["stepper"], T=T)
    TAD.set_c_(prob, c["c0"])
    ptf.stepforward_(prob, c["nsteps"])
    TAD.updatevars_(prob)
    return prob.vars.c, prob
