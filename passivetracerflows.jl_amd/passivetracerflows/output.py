"""Diagnostics and output: FourierFlows ``Diagnostic``/``increment!``/``Output``/
``saveproblem``/``saveoutput`` as PassiveTracerFlows uses them
(examples/onedim_gaussiandiffusion.jl:66-95), with a GPU-resident fast path.

``stepforward_(prob, diags, nsteps)`` keeps ``mass`` and ``variance``
diagnostics on the GPU (``ptf_step_record``) and copies them to the host once
per call; other diagnostic functions run on the host after every step, as in
FourierFlows.  ``record_`` also keeps snapshots of c on the GPU.

Output files: JLD2 (HDF5) cannot be written here, so an ``Output`` path is a
directory that holds one ``.npy`` file per JLD2 dataset under the same group
names (``grid/x``, ``snapshots/t/50``, ``snapshots/concentration/50`` ...).
``jldopen`` reads it back with JLD2-style keys.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

from ._lib import call


class Diagnostic:
    """Diagnostic(calc, prob; freq=1, nsteps=100, ndata=ceil((nsteps+1)/freq)):
    records calc(prob) at construction and then whenever clock.step % freq == 0
    after a step (increment_)."""

    def __init__(self, calc, prob, freq: int = 1, nsteps: int = 100, ndata: int | None = None):
        if freq < 1:
            raise ValueError("freq must be >= 1")
        self.calc = calc
        self.prob = prob
        self.freq = int(freq)
        self.ndata = int(ndata) if ndata is not None else int(math.ceil((nsteps + 1) / freq))
        self.data = [calc(prob)]
        self.t = [prob.clock.t]
        self.steps = [prob.clock.step]
        self.i = 1

    def _append(self, t, step, value):
        self.t.append(t)
        self.steps.append(step)
        self.data.append(value)
        self.i += 1
        self.ndata = max(self.ndata, self.i)

    @property
    def value(self):
        return self.data[self.i - 1]


def _as_list(diags):
    return list(diags) if isinstance(diags, (list, tuple)) else [diags]


def increment_(diags):
    """increment!(diags): record every diagnostic whose freq divides clock.step."""
    for d in _as_list(diags):
        if d.prob.clock.step % d.freq == 0:
            d._append(d.prob.clock.t, d.prob.clock.step, d.calc(d.prob))


def _device_kind(calc):
    from . import traceradvectiondiffusion as tad
    if calc is tad.mass:
        return "mass"
    if calc is tad.variance:
        return "variance"
    return None


def _step_record(prob, nsteps, every, want_c):
    """One ptf_step_record call; returns (t, steps, mass, variance, c or None)."""
    s0 = prob.clock.step
    nrec = (s0 + nsteps) // every - s0 // every
    cap = max(nrec, 1)
    t = np.zeros(cap)
    st = np.zeros(cap, np.int64)
    m = np.zeros(cap)
    v = np.zeros(cap)
    c = np.zeros((cap,) + tuple(prob.grid.shape)) if want_c else None
    n = C.c_int64()
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    call("ptf_step_record", prob._h, int(nsteps), int(every), cap, C.byref(n), dp(t),
         st.ctypes.data_as(C.POINTER(C.c_int64)), dp(m), dp(v), dp(c) if want_c else None)
    prob._sync_clock()
    k = n.value
    return t[:k], st[:k], m[:k], v[:k], (c[:k] if want_c else None)


def stepforward_diags_(prob, diags, nsteps: int):
    """stepforward!(prob, diags, nsteps): nsteps x (stepforward!(prob); increment!(diags))."""
    from . import fourierflows as ff
    from . import traceradvectiondiffusion as tad
    diags = _as_list(diags)
    nsteps = int(nsteps)
    if nsteps < 0:
        raise ValueError("nsteps must be >= 0")
    kinds = [_device_kind(d.calc) for d in diags]
    freqs = {d.freq for d in diags}
    device = (all(k is not None for k in kinds) and all(d.prob is prob for d in diags)
              and len(freqs) == 1 and not tad._is_host_time_varying(prob.params))
    if not device:
        for _ in range(nsteps):
            ff.stepforward_(prob, 1)
            increment_(diags)
        return None
    t, st, m, v, _ = _step_record(prob, nsteps, freqs.pop(), False)
    for d, k in zip(diags, kinds):
        vals = m if k == "mass" else v
        for i in range(len(t)):
            d._append(float(t[i]), int(st[i]), float(vals[i]))
    return None


def record_(prob, nsteps: int, every: int, concentration: bool = True):
    """Advance nsteps steps keeping, after every step whose clock.step is a multiple of
    `every`, t, step, mass, variance and (concentration=True) the field c on the GPU;
    returns them as numpy arrays, copied once at the end."""
    from . import fourierflows as ff
    from . import traceradvectiondiffusion as tad
    if tad._is_host_time_varying(prob.params):
        out = dict(t=[], step=[], mass=[], variance=[], c=[] if concentration else None)
        for _ in range(int(nsteps)):
            ff.stepforward_(prob, 1)
            if prob.clock.step % every == 0:
                out["t"].append(prob.clock.t)
                out["step"].append(prob.clock.step)
                out["mass"].append(tad.mass(prob))
                out["variance"].append(tad.variance(prob))
                if concentration:
                    tad.updatevars_(prob)
                    out["c"].append(prob.vars.c.copy())
        return {k: (np.asarray(v) if v is not None else None) for k, v in out.items()}
    t, st, m, v, c = _step_record(prob, int(nsteps), int(every), concentration)
    return dict(t=t, step=st, mass=m, variance=v, c=c)


# ---------------------------------------------------------------------------
# Output
# ---------------------------------------------------------------------------
def _uniquepath(path):
    """FourierFlows uniquepath: append _1, _2, ... before the extension while the path exists."""
    if not os.path.exists(path):
        return path
    root, ext = os.path.splitext(path)
    n = 1
    while os.path.exists(f"{root}_{n}{ext}"):
        n += 1
    return f"{root}_{n}{ext}"


def _write(path, key, value):
    f = os.path.join(path, *key.split("/")) + ".npy"
    os.makedirs(os.path.dirname(f), exist_ok=True)
    np.save(f, np.asarray(value), allow_pickle=False)


class Output:
    """Output(prob, filename, (name, func), ...): saveoutput writes func(prob) under
    snapshots/<name>/<step> and clock.t under snapshots/t/<step>."""

    def __init__(self, prob, filename, *fieldtuples):
        self.prob = prob
        self.path = _uniquepath(filename)
        self.fields = {}
        for ft in fieldtuples:
            name, func = ft
            self.fields[str(name)] = func
        os.makedirs(self.path, exist_ok=True)


def saveproblem(out: Output):
    """saveproblem(output): grid, clock.dt, eqn.L and the scalar parameters."""
    p = out.prob
    g = p.grid
    for d, (x, L, n) in enumerate((("x", "Lx", "nx"), ("y", "Ly", "ny"), ("z", "Lz", "nz"))):
        if d >= g.ndim:
            break
        _write(out.path, f"grid/{x}", getattr(g, x))
        _write(out.path, f"grid/{L}", float(getattr(g, L)))
        _write(out.path, f"grid/{n}", int(getattr(g, n)))
    _write(out.path, "clock/dt", float(p.clock.dt))
    L = p.eqn.L
    _write(out.path, "eqn/L", np.asarray(L))
    for k, v in vars(p.params).items():
        if isinstance(v, (int, float)) and not isinstance(v, bool):
            _write(out.path, f"params/{k}", v)


def saveoutput(out: Output):
    """saveoutput(output): one snapshot of every field at the current clock."""
    step = out.prob.clock.step
    _write(out.path, f"snapshots/t/{step}", float(out.prob.clock.t))
    for name, func in out.fields.items():
        _write(out.path, f"snapshots/{name}/{step}", np.asarray(func(out.prob)))


def save_records(out: Output, records, name: str = "concentration"):
    """Write the snapshots returned by record_ as saveoutput would have, one per step."""
    for i, step in enumerate(records["step"]):
        _write(out.path, f"snapshots/t/{int(step)}", float(records["t"][i]))
        if records.get("c") is not None:
            _write(out.path, f"snapshots/{name}/{int(step)}", records["c"][i])


class jldopen:
    """Read an Output directory with JLD2-style keys: f["snapshots/t/50"], f.keys("snapshots/t")."""

    def __init__(self, path):
        self.path = path

    def __getitem__(self, key):
        a = np.load(os.path.join(self.path, *key.split("/")) + ".npy", allow_pickle=False)
        return a.item() if a.ndim == 0 else a

    def keys(self, group=""):
        d = os.path.join(self.path, *group.split("/")) if group else self.path
        out = []
        for e in sorted(os.listdir(d)):
            out.append(e[:-4] if e.endswith(".npy") else e)
        return out

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False
