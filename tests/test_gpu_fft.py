"""GPU parity of the rfft plan (mul!/ldiv!) and of calcN! against the oracle.

Tolerances: FP64 relative L2 <= 1e-13 (transform round-off of both sides is
~1e-16 log2 N); FP32 <= 2e-6.
"""
import numpy as np
import pytest

import passivetracerflows as ptf
from oracle import tad as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


SHAPES = [(1, (16,)), (1, (128,)), (1, (4096,)), (2, (16, 8)), (2, (64, 32)), (2, (256, 256)),
          (2, (2048, 64)), (2, (32, 2048)), (3, (16, 8, 32)), (3, (64, 64, 64)), (3, (32, 256, 16))]


def _grid(dev, ndim, n, T=np.float64):
    if ndim == 1:
        return ptf.OneDGrid(dev, nx=n[0], T=T)
    if ndim == 2:
        return ptf.TwoDGrid(dev, nx=n[0], ny=n[1], T=T)
    return ptf.ThreeDGrid(dev, nx=n[0], ny=n[1], nz=n[2], T=T)


@pytest.mark.parametrize("ndim,n", SHAPES)
@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_rfft_irfft_vs_pocketfft(dev, ndim, n, T):
    print(f"rfft/irfft ndim={ndim} n={n} T={T.__name__}", flush=True)
    g = _grid(dev, ndim, n, T)
    rng = np.random.default_rng(1)
    c = rng.standard_normal(g.shape)
    ref = np.fft.rfftn(c, axes=tuple(range(ndim)))
    ct = {np.float64: np.complex128, np.float32: np.complex64}[T]
    din = ptf.DeviceArray.from_host(c.astype(T))
    dout = ptf.DeviceArray(g.spec_shape, ct)
    ptf.mul_(dout, g.rfftplan, din)
    got = dout.to_host()
    tol = 1e-13 if T is np.float64 else 2e-6
    assert rel(got, ref) < tol
    # input preserved
    assert np.array_equal(din.to_host(), c.astype(T))
    # ldiv! on a non-Hermitian spectrum: junk imaginary parts in the kx=0 and
    # kx=nx/2 columns are ignored after the y/z inverse passes
    sp = rng.standard_normal(g.spec_shape) + 1j * rng.standard_normal(g.spec_shape)
    dsp = ptf.DeviceArray.from_host(sp.astype(ct))
    dphys = ptf.DeviceArray(g.shape, T)
    ptf.ldiv_(dphys, g.rfftplan, dsp)
    back = dphys.to_host()
    refb = np.fft.irfftn(sp, s=g.shape, axes=tuple(range(ndim)))
    assert rel(back, refb) < tol
    assert np.array_equal(dsp.to_host(), sp.astype(ct))


def _random_problem(dev, ndim, n, T=np.float64, seed=0, kappa=0.0):
    rng = np.random.default_rng(seed)
    if ndim == 1:
        g = O.Grid(1, n[0])
        flows = (rng.standard_normal(g.shape),)
        prob = ptf.TracerAdvectionDiffusion.Problem(
            dev, ptf.OneDAdvectingFlow(u=lambda x: flows[0]), nx=n[0], κ=kappa, T=T)
    elif ndim == 2:
        g = O.Grid(2, n[0], ny=n[1])
        flows = (rng.standard_normal(g.shape), rng.standard_normal(g.shape))
        prob = ptf.TracerAdvectionDiffusion.Problem(
            dev, ptf.TwoDAdvectingFlow(u=lambda x, y: flows[0], v=lambda x, y: flows[1]),
            nx=n[0], ny=n[1], κ=kappa, T=T)
    else:
        g = O.Grid(3, n[0], ny=n[1], nz=n[2])
        flows = tuple(rng.standard_normal(g.shape) for _ in range(3))
        prob = ptf.TracerAdvectionDiffusion.Problem(
            dev, ptf.ThreeDAdvectingFlow(u=lambda x, y, z: flows[0], v=lambda x, y, z: flows[1],
                                         w=lambda x, y, z: flows[2]),
            nx=n[0], ny=n[1], nz=n[2], κ=kappa, T=T)
    return g, flows, prob, rng


@pytest.mark.parametrize("ndim,n", [(1, (16,)), (1, (256,)), (2, (32, 16)), (2, (128, 128)),
                                    (2, (512, 64)), (3, (16, 16, 16)), (3, (64, 32, 128))])
def test_calcN_vs_oracle(dev, ndim, n):
    print(f"calcN ndim={ndim} n={n}", flush=True)
    g, flows, prob, rng = _random_problem(dev, ndim, n)
    c = rng.standard_normal(g.shape)
    sol = O.rfft(g, c)
    # a non-Hermitian perturbation exercises the Nyquist/DC conventions too
    sol = sol + 1e-3 * (rng.standard_normal(sol.shape) + 1j * rng.standard_normal(sol.shape))
    ref = O.calcN(g, sol, flows)
    N = np.empty_like(sol)
    pr = prob
    ptf.TracerAdvectionDiffusion.calcN_(N, sol, 0.0, pr.clock, pr.vars, pr.params, pr.grid)
    assert rel(N, ref) < 1e-13


def test_linear_operator_vs_oracle(dev):
    g = O.Grid(3, 16, ny=8, nz=32, Lx=3.0, Ly=5.0, Lz=7.0)
    pr = ptf.TracerAdvectionDiffusion.Problem(dev, ptf.ThreeDAdvectingFlow(), nx=16, ny=8, nz=32,
                                              Lx=3.0, Ly=5.0, Lz=7.0, κ=0.3, η=0.2, ι=0.1)
    import ctypes as C
    L = np.empty(g.spec_shape)
    ptf._lib.call("ptf_get_L", pr._h, L.ctypes.data_as(C.POINTER(C.c_double)))
    ref = O.linear_operator(g, 0.3, 0.2, 0.1)
    assert rel(L, ref) < 1e-15


# layered (batched) plans: the (nx, ny, nlayers) transforms over dims 1:2 of the
# multi-layer tracer (S:678-679, S:865-866), and batched 1D/3D fields
LAYERED = [(2, (64, 32), 2), (2, (128, 128), 5), (1, (256,), 7), (3, (16, 32, 8), 3)]


@pytest.mark.parametrize("ndim,n,nl", LAYERED)
@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_layered_plan_vs_pocketfft(dev, ndim, n, nl, T):
    print(f"layered ndim={ndim} n={n} nlayers={nl} T={T.__name__}", flush=True)
    g = _grid(dev, ndim, n, T)
    plan = ptf.FourierFlows.RFFTPlan(g, nlayers=nl)
    rng = np.random.default_rng(2)
    c = rng.standard_normal((nl,) + g.shape)
    axes = tuple(range(1, ndim + 1))
    ref = np.fft.rfftn(c, axes=axes)
    ct = {np.float64: np.complex128, np.float32: np.complex64}[T]
    din = ptf.DeviceArray.from_host(c.astype(T))
    dout = ptf.DeviceArray((nl,) + g.spec_shape, ct)
    ptf.mul_(dout, plan, din)
    tol = 1e-13 if T is np.float64 else 2e-6
    got = dout.to_host()
    for layer in range(nl):          # each layer on its own, none mixed into another
        assert rel(got[layer], ref[layer]) < tol
    back = ptf.DeviceArray((nl,) + g.shape, T)
    ptf.ldiv_(back, plan, dout)
    assert rel(back.to_host(), c) < tol
    # dealias! on every layer
    ptf.dealias_(dout, g, plan)
    og = O.Grid(ndim, n[0]) if ndim == 1 else (O.Grid(2, n[0], ny=n[1]) if ndim == 2
                                              else O.Grid(3, n[0], ny=n[1], nz=n[2]))
    d = dout.to_host()
    for layer in range(nl):
        assert np.array_equal(d[layer], O.dealias(og, got[layer]))
