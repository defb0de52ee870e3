"""Seeded synthetic inputs for the five measured TracerAdvectionDiffusion workloads.

Each ``make_*`` returns a plain dict of numpy FP64 arrays and scalars; the
same dict drives the HIP problems and the CPU checks.  Physical arrays are
C-ordered ``(nz, ny, nx)`` (x fastest), as the library and the Julia
column-major layout expect.

* ``cellular_2d``    -- 2D steady cellular flow psi = (U/k) sin(kx) sin(ky), Gaussian blob.
* ``abc_3d``         -- 3D steady ABC flow, random-phase |c_k|^2 ~ k^-2 (1<=|k|<=8) initial field.
* ``gaussian_1d``    -- 1D pure diffusion of a Gaussian.
* ``turbulent_2d``   -- 2D random-phase streamfunction with E(k) ~ k^3 exp(-(k/8)^2) whose
                        phases advance at omega = |k|^0.5; hyperdiffusion order 4; ETDRK4.
* ``ensemble_2d``    -- 512 cellular-flow problems of sizes 64^2..512^2 in ratio 8:4:2:1.
"""
from __future__ import annotations

import numpy as np

TWO_PI = 2 * np.pi


def _points(n: int, L: float = TWO_PI) -> np.ndarray:
    start = np.longdouble(-L / 2)
    step = np.longdouble(L / n)
    return (start + np.arange(n, dtype=np.longdouble) * step).astype(np.float64)


def grid_xyz(ndim: int, n):
    """Coordinate arrays broadcast to the physical shape (x fastest)."""
    ns = [n] * ndim if np.isscalar(n) else list(n)
    axes = [_points(k) for k in ns]
    if ndim == 1:
        return (axes[0],)
    if ndim == 2:
        Y, X = np.meshgrid(axes[1], axes[0], indexing="ij")
        return X, Y
    Z, Y, X = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return X, Y, Z


def _wrap(d):
    return (d + np.pi) % TWO_PI - np.pi


def gaussian_1d(nx=128, nsteps=1000):
    (x,) = grid_xyz(1, nx)
    return dict(name="1D Gaussian diffusion", ndim=1, n=(nx, 1, 1), L=(TWO_PI,) * 3,
                kappa=0.01, eta=0.01, iota=0.01, kappah=0.0, nkappah=0,
                dt=0.01, nsteps=nsteps, stepper="RK4",
                flow=(np.zeros(nx),), c0=np.exp(-x ** 2 / (2 * 0.1 ** 2)))


def cellular_flow(X, Y, U=1.0, k=1.0):
    """u = -d(psi)/dy, v = d(psi)/dx for psi = (U/k) sin(kx) sin(ky)."""
    return -U * np.sin(k * X) * np.cos(k * Y), U * np.cos(k * X) * np.sin(k * Y)


def cellular_2d(n=256, nsteps=2000, kappa=2e-3, dt=0.01):
    X, Y = grid_xyz(2, n)
    u, v = cellular_flow(X, Y)
    c0 = np.exp(-((X - 0.5) ** 2 + Y ** 2) / (2 * 0.2 ** 2))
    return dict(name=f"2D steady cellular flow {n}^2", ndim=2, n=(n, n, 1), L=(TWO_PI,) * 3,
                kappa=kappa, eta=kappa, iota=kappa, kappah=0.0, nkappah=0,
                dt=dt, nsteps=nsteps, stepper="RK4", flow=(u, v), c0=c0)


def random_field_3d(n, seed=3, kmin=1.0, kmax=8.0):
    """Real field whose half-spectrum is |k|^-1 e^{i theta} on 1<=|k|<=8, unit variance."""
    rng = np.random.Generator(np.random.PCG64(seed))
    kx = np.arange(n // 2 + 1)[None, None, :]
    kf = np.fft.fftfreq(n, 1.0 / n)
    ky = kf[None, :, None]
    kz = kf[:, None, None]
    K = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    shell = (K >= kmin) & (K <= kmax)
    theta = rng.uniform(0.0, TWO_PI, size=K.shape)
    ch = np.where(shell, np.exp(1j * theta) / np.where(K > 0, K, 1.0), 0.0)
    c = np.fft.irfftn(ch, s=(n, n, n), axes=(0, 1, 2))
    c = np.fft.irfftn(np.fft.rfftn(c, axes=(0, 1, 2)), s=(n, n, n), axes=(0, 1, 2))
    c -= c.mean()
    c /= c.std()
    return c


def abc_3d(n=256, nsteps=500, A=1.0, B=1.0, C=1.0, seed=3, dt=5e-3, kappa=1e-3):
    X, Y, Z = grid_xyz(3, n)
    u = A * np.sin(Z) + C * np.cos(Y)
    v = B * np.sin(X) + A * np.cos(Z)
    w = C * np.sin(Y) + B * np.cos(X)
    return dict(name=f"3D steady ABC flow {n}^3", ndim=3, n=(n, n, n), L=(TWO_PI,) * 3,
                kappa=kappa, eta=kappa, iota=kappa, kappah=0.0, nkappah=0,
                dt=dt, nsteps=nsteps, stepper="RK4", flow=(u, v, w),
                c0=random_field_3d(n, seed))


def turbulent_psih(n, seed=4, kpeak=8.0):
    """Half spectrum psi_0 (ny, nx/2+1): |psi|^2 ~ E(k)/k^3, E ~ k^3 exp(-(k/kpeak)^2),
    random phases; scaled so that sqrt(mean(u^2 + v^2)) = 1.  Returns (psih0, omega)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    kx = np.arange(n // 2 + 1, dtype=np.float64)[None, :]
    ky = np.fft.fftfreq(n, 1.0 / n)[:, None]
    K = np.sqrt(kx ** 2 + ky ** 2)
    Ks = np.where(K > 0, K, 1.0)
    E = Ks ** 3 * np.exp(-(Ks / kpeak) ** 2)
    amp = np.where(K > 0, np.sqrt(E / Ks ** 3), 0.0)
    theta = rng.uniform(0.0, TWO_PI, size=K.shape)
    psih = amp * np.exp(1j * theta)
    # Hermitian projection so that psi_0 is real
    psi = np.fft.irfft2(psih, s=(n, n))
    psih = np.fft.rfft2(psi)
    u = np.fft.irfft2(-1j * ky * psih, s=(n, n))
    v = np.fft.irfft2(1j * kx * psih, s=(n, n))
    rms = np.sqrt(np.mean(u ** 2 + v ** 2))
    psih /= rms
    omega = np.sqrt(K)
    return psih, omega


def turbulent_2d(n=2048, nsteps=1000, seed=4, dt=2e-3):
    X, Y = grid_xyz(2, n)
    psih0, omega = turbulent_psih(n, seed)
    rng = np.random.Generator(np.random.PCG64(seed + 100))
    c0 = np.sin(X) + 0.01 * rng.standard_normal(X.shape)
    return dict(name=f"2D unsteady random-phase flow {n}^2", ndim=2, n=(n, n, 1), L=(TWO_PI,) * 3,
                kappa=0.0, eta=0.0, iota=0.0, kappah=1e-12, nkappah=4,
                dt=dt, nsteps=nsteps, stepper="ETDRK4", psih0=psih0, omega=omega, c0=c0)


def spectral_flow_at(psih0, omega, t, n):
    """u, v at time t for psi_hat(t) = psi_0 exp(-i omega t) on the stored half spectrum."""
    kx = np.arange(n // 2 + 1, dtype=np.float64)[None, :]
    ky = np.fft.fftfreq(n, 1.0 / n)[:, None]
    ph = psih0 * np.exp(-1j * omega * t)
    return (np.fft.irfft2(-1j * ky * ph, s=(n, n)), np.fft.irfft2(1j * kx * ph, s=(n, n)))


ENSEMBLE_SIZES = (64, 128, 256, 512)
ENSEMBLE_WEIGHTS = (8, 4, 2, 1)


def ensemble_counts(total=512):
    w = np.array(ENSEMBLE_WEIGHTS, dtype=np.float64)
    counts = [int(round(total * x / w.sum())) for x in w]
    counts[0] += total - sum(counts)
    return counts


def ensemble_member(n, seed):
    """One cellular-flow member with seeded kappa, k, U, Gaussian centre and width."""
    rng = np.random.Generator(np.random.PCG64(seed))
    kappa = float(np.exp(rng.uniform(np.log(1e-4), np.log(1e-2))))
    k = float(rng.choice([1, 2, 4]))
    U = float(rng.uniform(0.5, 2.0))
    x0, y0 = rng.uniform(-np.pi, np.pi, size=2)
    sigma = float(rng.uniform(0.1, 0.4))
    X, Y = grid_xyz(2, n)
    u, v = cellular_flow(X, Y, U, k)
    c0 = np.exp(-(_wrap(X - x0) ** 2 + _wrap(Y - y0) ** 2) / (2 * sigma ** 2))
    return dict(n=n, kappa=kappa, k=k, U=U, sigma=sigma, centre=(float(x0), float(y0)),
                flow=(u, v), c0=c0)


def ensemble_2d(total=512, seed=5, dt=5e-3, nsteps=1000, counts=None):
    """Member descriptors (lazy: arrays are built by ``ensemble_member``)."""
    counts = ensemble_counts(total) if counts is None else list(counts)
    members = []
    idx = 0
    for n, c in zip(ENSEMBLE_SIZES, counts):
        for _ in range(c):
            members.append(dict(n=n, seed=seed * 100000 + idx))
            idx += 1
    return dict(name=f"2D cellular-flow ensemble of {len(members)}", members=members, dt=dt,
                nsteps=nsteps, stepper="RK4", counts=counts)


RK4_REAL_LIMIT = 2.785293563405282  # RK4 stability interval on the negative real axis


def spectral_tail(solh, n):
    """Fraction of the L2 norm of a 2D half spectrum held by modes with |kx| or |ky| > n/3
    (the band the 2/3 rule would remove); a resolved field keeps this small."""
    kx = np.arange(n // 2 + 1)[None, :]
    ky = np.abs(np.fft.fftfreq(n, 1.0 / n))[:, None]
    mask = np.broadcast_to((kx > n / 3) | (ky > n / 3), solh.shape)
    tot = np.linalg.norm(solh)
    return float(np.linalg.norm(solh[mask]) / tot) if tot > 0 else 0.0


def member_class(m, dt, sol64, tail_limit=1e-2):
    """'diffusion_unstable': kappa |k_max|^2 dt beyond RK4's real-axis limit (the FP64 run
    overflows, as the reference algorithm does); 'under_resolved': the FP64 state at the end
    holds more than 1 % of its norm in the |k| > n/3 band (grid-scale noise of the
    non-dealiased scheme); otherwise 'resolved'."""
    n = m["n"]
    if m["kappa"] * 2 * (n / 2) ** 2 * dt > RK4_REAL_LIMIT:
        return "diffusion_unstable"
    with np.errstate(over="ignore", invalid="ignore"):
        tail = spectral_tail(sol64, n) if np.all(np.isfinite(sol64)) else float("nan")
    if not tail <= tail_limit:   # NaN (overflowed norm) counts as under-resolved
        return "under_resolved"
    return "resolved"
