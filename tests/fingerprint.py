"""Compact fingerprints of a spectral state, for parity checks of long runs.

A full 256^3 or 2048^2 spectral field is too large to commit as a fixture, so
the full-length oracle runs are stored as a fingerprint that a plausible
error cannot pass unnoticed:

* ``low``   -- every mode with |kx|, |ky|, |kz| <= K (the energetic scales);
* ``norm``  -- the L2 norm of the whole half spectrum;
* ``mass``  -- Re of the k = 0 mode (= sum of c over the grid); its error is
               measured against the norm of the state (the share of the relative
               L2 error that the k = 0 mode carries);
* ``var``   -- the variance of c by Parseval;
* ``proj``  -- inner products <w_i, sol> with seeded complex Gaussian fields w_i
               over the whole half spectrum.  For a difference d = sol - ref,
               <w_i, d> is a zero-mean Gaussian with standard deviation
               sqrt(2) ||d||, so |<w_i, sol> - <w_i, ref>| / ||ref|| estimates
               the relative L2 error of the whole field.
"""
from __future__ import annotations

import numpy as np

NPROJ = 4


def _low_index(shape, K):
    """Index arrays of the |k| <= K box of a half spectrum of the given shape."""
    nkr = shape[-1]
    idx = [np.arange(min(K, nkr - 1) + 1)]
    for n in shape[:-1][::-1]:
        k = min(K, n // 2 - 1)     # small grids: the box stops short of the Nyquist row
        idx.append(np.concatenate([np.arange(0, k + 1), np.arange(n - k, n)]))
    # numpy axis order is (z, y, x)
    return np.ix_(*idx[::-1])


def projection_weights(shape, seed, i):
    rng = np.random.Generator(np.random.PCG64([seed, i]))
    w = rng.standard_normal(shape + (2,))
    return w[..., 0] + 1j * w[..., 1]


def variance(sol, npoints):
    w = np.full(sol.shape[-1], 2.0)
    w[0] = 1.0
    w[-1] = 1.0
    p = (np.abs(sol) ** 2) * w
    tot = p.sum() - abs(sol.reshape(-1)[0]) ** 2
    return float(tot / npoints ** 2)


def fingerprint(sol, npoints, K=16, seed=2024):
    sol = np.asarray(sol, dtype=np.complex128)
    proj = np.array([np.vdot(projection_weights(sol.shape, seed, i), sol) for i in range(NPROJ)])
    return dict(low=sol[_low_index(sol.shape, K)].copy(), norm=float(np.linalg.norm(sol)),
                mass=float(sol.reshape(-1)[0].real), var=variance(sol, npoints), proj=proj,
                K=K, seed=seed)


def compare(sol, ref: dict, npoints):
    """Relative errors of every fingerprint entry of ``sol`` against ``ref`` (a stored
    fingerprint): low-mode block (relative L2), norm, variance (relative), mass (relative
    to the norm of the state, since the mass of a zero-mean field is roundoff) and the
    largest projection error relative to the norm."""
    fp = fingerprint(sol, npoints, K=int(ref["K"]), seed=int(ref["seed"]))
    nref = float(ref["norm"])
    low_ref = np.asarray(ref["low"])
    return dict(
        low=float(np.linalg.norm(fp["low"] - low_ref) / np.linalg.norm(low_ref)),
        norm=abs(fp["norm"] - nref) / nref,
        var=abs(fp["var"] - float(ref["var"])) / abs(float(ref["var"])),
        mass=abs(fp["mass"] - float(ref["mass"])) / nref,
        proj=float(np.max(np.abs(fp["proj"] - np.asarray(ref["proj"])))) / (np.sqrt(2) * nref),
    )
