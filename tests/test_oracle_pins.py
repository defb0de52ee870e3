"""Pins of the oracle restatement (CPU only).

Each check is chosen so that a plausible slip (a dropped term, a wrong sign,
index or normalisation) fails it without retyping the oracle's formulas:

* transforms       -- brute-force DFT in long double; non-Hermitian C2R convention
* grid             -- exact gridpoints, Nyquist signs of rfftfreq/fftfreq
* L                -- explicit values at chosen modes; 0^0 = 1
* calcN            -- constant flow gives -i (k.U) c_hat; linearity; zero mean tendency
* RK4              -- closed-form amplification R(z)^n, z = dt (L - i k.U), per mode;
                      time-varying uniform flow frozen at clock.t: prod R(z(t_m))
* ETDRK4           -- no flow: exp(L dt n) exactly; constant flow: 4th-order convergence
* phi-coefficients -- mpmath direct evaluation at 50 digits
* reference tests  -- test/runtests.jl:26-54 analytic solutions at their tolerances
"""
import numpy as np
import pytest

from oracle import tad as O
from reference_cases import NAMES, case, isapprox, run_oracle


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def brute_dft(x, axes_sign=-1):
    """Full complex DFT over all axes in long double."""
    x = np.asarray(x, dtype=np.clongdouble)
    out = x
    for ax in range(x.ndim):
        n = x.shape[ax]
        k = np.arange(n, dtype=np.longdouble)
        ang = (2 * np.pi * np.outer(k, k) / n).astype(np.longdouble)
        W = np.cos(ang) + axes_sign * 1j * np.sin(ang)
        out = np.moveaxis(np.tensordot(W, np.moveaxis(out, ax, 0), axes=(1, 0)), 0, ax)
    return out


@pytest.mark.parametrize("shape", [(16,), (8, 12), (4, 6, 8)])
def test_rfft_vs_bruteforce(shape):
    rng = np.random.default_rng(0)
    nd = len(shape)
    g = O.Grid(nd, shape[-1], ny=shape[-2] if nd >= 2 else None, nz=shape[0] if nd == 3 else None)
    c = rng.standard_normal(shape)
    full = brute_dft(c)
    half = full[..., : shape[-1] // 2 + 1]
    assert rel(O.rfft(g, c), half.astype(np.complex128)) < 1e-15


@pytest.mark.parametrize("shape", [(16,), (8, 12), (4, 6, 8)])
def test_irfft_nonhermitian_convention(shape):
    """C2R = complex inverse over the non-x axes, then along x a Hermitian extension that
    keeps only the real part of the kx = 0 and kx = nx/2 entries."""
    rng = np.random.default_rng(1)
    nd = len(shape)
    g = O.Grid(nd, shape[-1], ny=shape[-2] if nd >= 2 else None, nz=shape[0] if nd == 3 else None)
    nx = shape[-1]
    M = nx // 2
    sp = rng.standard_normal(g.spec_shape) + 1j * rng.standard_normal(g.spec_shape)
    # complex inverse over the leading axes (long double), then Hermitian x-extension
    lead = sp.astype(np.clongdouble)
    for ax in range(nd - 1):
        n = shape[ax]
        k = np.arange(n, dtype=np.longdouble)
        ang = (2 * np.pi * np.outer(k, k) / n).astype(np.longdouble)
        Wi = (np.cos(ang) + 1j * np.sin(ang)) / n
        lead = np.moveaxis(np.tensordot(Wi, np.moveaxis(lead, ax, 0), axes=(1, 0)), 0, ax)
    ext = np.zeros(shape, dtype=np.clongdouble)
    ext[..., : M + 1] = lead
    ext[..., 0] = lead[..., 0].real
    ext[..., M] = lead[..., M].real
    for k in range(1, M):
        ext[..., nx - k] = np.conj(lead[..., k])
    kk = np.arange(nx, dtype=np.longdouble)
    ang = (2 * np.pi * np.outer(kk, kk) / nx).astype(np.longdouble)
    Wx = (np.cos(ang) + 1j * np.sin(ang)) / nx
    ref = np.tensordot(ext, Wx, axes=([nd - 1], [0]))
    assert np.abs(ref.imag).max() < 1e-15
    assert rel(O.irfft(g, sp), ref.real.astype(np.float64)) < 1e-14


def test_gridpoints_and_wavenumbers():
    g = O.Grid(3, 8, ny=6 * 2, nz=4)
    assert g.x[0] == -np.pi and g.x[4] == 0.0
    assert np.allclose(np.diff(g.x), 2 * np.pi / 8, rtol=0, atol=1e-15)
    assert list(O.rfftfreq_k(8, 2 * np.pi)) == [0, 1, 2, 3, 4]          # Nyquist +n/2
    assert list(O.fftfreq_k(8, 2 * np.pi)) == [0, 1, 2, 3, -4, -3, -2, -1]  # Nyquist -n/2
    assert np.allclose(O.rfftfreq_k(8, 4.0), np.arange(5) * 2 * np.pi / 4.0, rtol=1e-15)


def test_linear_operator_values():
    g = O.Grid(3, 8, Lx=2 * np.pi, ny=4, Ly=np.pi, nz=4, Lz=4 * np.pi)
    L = O.linear_operator(g, 0.1, 0.2, 0.3, kappah=0.01, nkappah=2)
    # mode (kx=3, ky=-2 -> index 2, kz=1): kr = 3, l = -2 * 2 = -4, m = 1 * 0.5
    kr, l, m = 3.0, -4.0, 0.5
    K2 = kr ** 2 + l ** 2 + m ** 2
    assert L[1, 2, 3] == pytest.approx(-0.1 * 9 - 0.2 * 16 - 0.3 * 0.25 - 0.01 * K2 ** 2, rel=1e-15)
    # 0^0 = 1: nkappah = 0 damps every mode, including k = 0
    L0 = O.linear_operator(O.Grid(2, 8), 0.0, 0.0, kappah=0.5, nkappah=0)
    assert np.all(L0 == -0.5)
    # 1D uses (kr^2)^n
    L1 = O.linear_operator(O.Grid(1, 8), 0.0, kappah=1.0, nkappah=3)
    assert L1[2] == -(4.0 ** 3)


def _bandlimited(g, rng):
    c = rng.standard_normal(g.shape)
    ch = O.rfft(g, c)
    mask = np.ones(g.spec_shape, bool)
    mask[..., -1] = False                       # kx = nx/2
    if g.ndim >= 2:
        mask[..., g.ny // 2, :] = False         # ky = -ny/2
    if g.ndim == 3:
        mask[g.nz // 2, :, :] = False           # kz = -nz/2
    return ch * mask


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_calcN_constant_flow(nd):
    rng = np.random.default_rng(2)
    g = O.Grid(nd, 16, ny=8 if nd >= 2 else None, nz=12 if nd == 3 else None)
    U = (0.3, -0.7, 1.1)[:nd]
    sol = _bandlimited(g, rng)
    flow = tuple(np.full(g.shape, u) for u in U)
    N = O.calcN(g, sol, flow)
    kdotU = g.kr * U[0] + (g.l * U[1] if nd >= 2 else 0) + (g.m * U[2] if nd == 3 else 0)
    assert rel(N, -1j * kdotU * sol) < 1e-14
    # linearity and zero flow
    sol2 = _bandlimited(g, rng)
    assert rel(O.calcN(g, 2 * sol - sol2, flow), 2 * N - O.calcN(g, sol2, flow)) < 1e-13
    assert np.abs(O.calcN(g, sol, tuple(np.zeros(g.shape) for _ in U))).max() == 0


def test_calcN_mass_divergence_free():
    """N(0) = -sum(u . grad c) = sum(c div u) = 0 for a divergence-free flow."""
    g = O.Grid(2, 32)
    X, Y = g.gridpoints()
    u, v = -np.sin(X) * np.cos(Y), np.cos(X) * np.sin(Y)
    rng = np.random.default_rng(3)
    N = O.calcN(g, _bandlimited(g, rng), (u, v))
    assert abs(N.reshape(-1)[0]) < 1e-12 * np.abs(N).max()


def R4(z):
    return 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_rk4_closed_form_constant_flow(nd):
    rng = np.random.default_rng(4)
    g = O.Grid(nd, 16, ny=8 if nd >= 2 else None, nz=8 if nd == 3 else None)
    U = (0.4, -0.3, 0.2)[:nd]
    dt, n, kappa = 0.05, 30, 0.02
    p = O.Problem(g, tuple(np.full(g.shape, u) for u in U), kappa=kappa, eta=kappa / 2,
                  iota=kappa / 3, dt=dt)
    sol0 = _bandlimited(g, rng)
    p.sol = sol0.copy()
    p.stepforward(n)
    kdotU = g.kr * U[0] + (g.l * U[1] if nd >= 2 else 0) + (g.m * U[2] if nd == 3 else 0)
    z = dt * (p.L - 1j * kdotU)
    assert rel(p.sol, R4(z) ** n * sol0) < 1e-13
    assert p.clock.step == n and abs(p.clock.t - n * dt) < 1e-14


def test_rk4_time_varying_frozen_at_clock_t():
    """U(t) uniform: each step uses U(clock.t) for all four stages."""
    rng = np.random.default_rng(5)
    g = O.Grid(1, 32)
    dt, n = 0.02, 25
    Uf = lambda t: 0.5 + 3.0 * t
    p = O.Problem(g, lambda t: (np.full(g.shape, Uf(t)),), kappa=0.01, dt=dt)
    sol0 = _bandlimited(g, rng)
    p.sol = sol0.copy()
    p.stepforward(n)
    amp = np.ones_like(sol0)
    for m in range(n):
        amp = amp * R4(dt * (p.L - 1j * g.kr * Uf(m * dt)))
    assert rel(p.sol, amp * sol0) < 1e-13
    # stage-time evaluation would differ measurably
    assert rel(p.sol, R4(dt * (p.L - 1j * g.kr * Uf(n * dt / 2))) ** n * sol0) > 1e-6


def test_etdrk4_no_flow_is_exponential():
    rng = np.random.default_rng(6)
    g = O.Grid(2, 16)
    dt, n = 0.1, 20
    p = O.Problem(g, (np.zeros(g.shape), np.zeros(g.shape)), kappa=0.3, eta=0.1, kappah=1e-3,
                  nkappah=2, dt=dt, stepper="ETDRK4")
    sol0 = _bandlimited(g, rng)
    p.sol = sol0.copy()
    p.stepforward(n)
    assert rel(p.sol, np.exp(p.L * dt * n) * sol0) < 1e-13


def test_etdrk4_fourth_order_constant_flow():
    """Constant flow: exact solution exp((L - i kU) t) c0; ETDRK4 error ~ dt^4."""
    rng = np.random.default_rng(7)
    g = O.Grid(1, 16)
    sol0 = _bandlimited(g, rng)
    U, kappa, T = 0.7, 0.05, 1.0
    errs = []
    for n in (20, 40, 80):
        dt = T / n
        p = O.Problem(g, (np.full(g.shape, U),), kappa=kappa, dt=dt, stepper="ETDRK4")
        p.sol = sol0.copy()
        p.stepforward(n)
        exact = np.exp((p.L - 1j * g.kr * U) * T) * sol0
        errs.append(rel(p.sol, exact))
    orders = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(3.7 < o < 4.5 for o in orders), (errs, orders)


def test_rk4_fourth_order_sanity():
    rng = np.random.default_rng(8)
    g = O.Grid(1, 16)
    sol0 = _bandlimited(g, rng)
    errs = []
    for n in (20, 40, 80):
        dt = 1.0 / n
        p = O.Problem(g, (np.full(g.shape, 0.7),), kappa=0.05, dt=dt)
        p.sol = sol0.copy()
        p.stepforward(n)
        errs.append(rel(p.sol, np.exp((p.L - 1j * g.kr * 0.7) * 1.0) * sol0))
    orders = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(3.7 < o < 4.5 for o in orders), (errs, orders)


def test_etd_coefficients_vs_mpmath():
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 50
    dt = 0.01
    Ls = np.array([0.0, -1e-6, -0.5, -3.0, -40.0, -250.0, -1e4, -3.8e10])
    E, E2, ze, al, be, ga = O.etd_coefficients(dt, Ls)
    for i, L in enumerate(Ls):
        z = mpmath.mpf(float(dt * L))  # the double the oracle forms
        if z == 0:
            ref = (dt / 2, dt / 6, dt / 6, dt / 6)
        else:
            ez = mpmath.e ** z
            ref = (dt * (mpmath.e ** (z / 2) - 1) / z,
                   dt * (-4 - z + ez * (4 - 3 * z + z ** 2)) / z ** 3,
                   dt * (2 + z + ez * (-2 + z)) / z ** 3,
                   dt * (-4 - 3 * z - z ** 2 + ez * (4 - z)) / z ** 3)
        for got, r in zip((ze[i], al[i], be[i], ga[i]), ref):
            r = float(r)
            assert abs(got - r) <= 1e-13 * abs(r) + 1e-300, (L, got, r)
        assert E[i] == pytest.approx(float(mpmath.e ** z), rel=1e-15, abs=1e-300)


def test_set_c_updatevars_mass_variance():
    rng = np.random.default_rng(9)
    g = O.Grid(3, 8, ny=4, nz=6)
    c = rng.standard_normal(g.shape)
    p = O.Problem(g, tuple(np.zeros(g.shape) for _ in range(3)))
    p.set_c(c)
    assert rel(p.c, c) < 1e-15
    assert O.mass(g, p.sol) == pytest.approx(c.sum(), rel=1e-13)
    assert O.variance(g, p.sol) == pytest.approx(c.var(), rel=1e-13)


@pytest.mark.parametrize("name", NAMES)
def test_reference_analytic_cases_oracle(name):
    c = case(name)
    assert isapprox(c["cfinal"], run_oracle(c), c["rtol"]), name


def test_unknown_stepper():
    with pytest.raises(ValueError):
        O.Problem(O.Grid(1, 8), (np.zeros(8),), stepper="AB3")


# ---------------------------------------------------------------------------
# filtered steppers and dealiasing ([FF] makefilter, Filtered*update!, dealias!)
# ---------------------------------------------------------------------------
def test_makefilter_closed_forms():
    """K = |(kr dx, l dy, m dz)|/pi is 2|i|/n for index i whatever L is; the filter is 1
    below innerK and equals tol at K = outerK (the x-Nyquist column for outerK = 1)."""
    g = O.Grid(2, 64)
    f = O.makefilter(g)
    Kx = 2 * np.arange(33) / 64
    decay = -np.log(1e-15) / 0.35 ** 4
    expect = np.where(Kx < 0.65, 1.0, np.exp(-decay * (Kx - 0.65) ** 4))
    assert np.allclose(f[0], expect, rtol=1e-13, atol=0)
    assert abs(f[0, 32] - 1e-15) <= 1e-27
    assert np.all(f[0, Kx < 0.65] == 1.0)
    assert np.all(np.diff(f[0]) <= 0)
    # independent of the domain lengths; even in l
    g2 = O.Grid(2, 64, Lx=3.0, Ly=5.0)
    assert np.allclose(O.makefilter(g2), f, rtol=1e-11, atol=1e-300)
    ky = np.fft.fftfreq(64, 1 / 64)
    for i in range(1, 32):
        j = int(np.where(ky == -ky[i])[0][0])
        assert np.allclose(f[i], f[j], rtol=1e-15, atol=0)
    # 3D: K adds the m direction; custom order/innerK/outerK/tol reach tol at outerK
    g3 = O.Grid(3, 16)
    f3 = O.makefilter(g3, order=6, innerK=0.5, outerK=0.75, tol=1e-10)
    # index 6 on every axis: K = sqrt(3) * 0.75 > outerK; index (0, 0, 6): K = 0.75
    assert abs(f3[0, 0, 6] - 1e-10) <= 1e-22
    assert f3[6, 6, 6] < 1e-10


def _zero_flow_problem(ndim, n, stepper, kappa, dt, seed, **fk):
    g = O.Grid(ndim, n)
    rng = np.random.default_rng(seed)
    flows = tuple(np.zeros(g.shape) for _ in range(ndim))
    p = O.Problem(g, flows, kappa=kappa, dt=dt, stepper=stepper, **fk)
    p.set_c(rng.standard_normal(g.shape))
    return g, p


@pytest.mark.parametrize("ndim,n", [(1, 64), (2, 32), (3, 16)])
def test_filtered_rk4_zero_flow_is_filter_times_rk4_polynomial(ndim, n):
    """u = 0: each FilteredRK4 step multiplies sol by filter * R(dt L), R the RK4 polynomial
    (the filter acts once per step, after the update)."""
    g, p = _zero_flow_problem(ndim, n, "FilteredRK4", 0.02, 0.1, 3)
    sol0 = p.sol.copy()
    p.stepforward(7)
    z = 0.1 * p.L
    R = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    expect = (O.makefilter(g) * R) ** 7 * sol0
    assert np.linalg.norm(p.sol - expect) <= 1e-13 * np.linalg.norm(expect)


def test_filtered_etdrk4_zero_flow_is_filter_times_exp():
    g, p = _zero_flow_problem(2, 32, "FilteredETDRK4", 0.02, 0.1, 4, order=6, innerK=0.5,
                              outerK=0.9, tol=1e-12)
    sol0 = p.sol.copy()
    p.stepforward(5)
    expect = (O.makefilter(g, order=6, innerK=0.5, outerK=0.9, tol=1e-12)
              * np.exp(0.1 * p.L)) ** 5 * sol0
    assert np.linalg.norm(p.sol - expect) <= 1e-13 * np.linalg.norm(expect)


@pytest.mark.parametrize("nk", [8, 16, 64, 128, 256, 1024])
def test_aliased_ranges_two_thirds_rule(nk):
    """The aliased set of a full axis is one contiguous block around the Nyquist index with
    all kept |k| < nk/3 and about a third of the modes; the half axis drops the same kr."""
    (lo, hi), (rlo, rhi) = O.aliased_ranges(nk, nk // 2 + 1)
    k = np.fft.fftfreq(nk, 1 / nk)
    al = np.zeros(nk, bool)
    al[lo:hi] = True
    assert al[nk // 2]                        # Nyquist is aliased
    assert np.all(np.abs(k[~al]) < nk / 3)
    assert abs(al.sum() - nk / 3) <= 2
    assert rlo == lo and rhi == nk // 2 + 1
    # f = 0: [FF] keeps only the Nyquist entries in the ranges and dealias! does nothing
    (zlo, zhi), (zr0, zr1) = O.aliased_ranges(nk, nk // 2 + 1, 0.0)
    assert (zlo, zhi) == (nk // 2, nk // 2 + 1) and (zr0, zr1) == (nk // 2, nk // 2 + 1)


def test_dealias_zeroes_exactly_the_aliased_band():
    g = O.Grid(3, 16, ny=32, nz=8)
    fh = np.ones(g.spec_shape, np.complex128)
    d = O.dealias(g, fh)
    (ylo, yhi), _ = O.aliased_ranges(32, 17)
    (zlo, zhi), _ = O.aliased_ranges(8, 5)
    _, (rlo, _) = O.aliased_ranges(16, 9)
    for iz in range(8):
        for iy in range(32):
            for ix in range(9):
                zero = (ix >= rlo) or (ylo <= iy < yhi) or (zlo <= iz < zhi)
                assert d[iz, iy, ix] == (0 if zero else 1)
    assert np.array_equal(O.dealias(g, fh, 0.0), fh)
