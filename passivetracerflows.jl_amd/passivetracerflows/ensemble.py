"""Ensembles of independent 2D tracer problems advanced by one graph-captured launch sequence.

Each member is a ``TracerAdvectionDiffusion.Problem(dev, TwoDAdvectingFlow(u, v);
nx, κ, η, dt, stepper="RK4")`` with a steady flow.  Members with the same grid
share one batched set of kernels; the per-step launches of all grid sizes are
captured once into a hipGraph (one branch per size) and replayed.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import call, dptr
from .fourierflows import variant_flags


def _gridpoints2(n, L=2 * np.pi):
    start = np.longdouble(-L / 2)
    step = np.longdouble(L / n)
    x = (start + np.arange(n, dtype=np.longdouble) * step).astype(np.float64)
    Y, X = np.meshgrid(x, x, indexing="ij")
    return X, Y


class Ensemble:
    def __init__(self, members, dt, precision="mixed", stepper="RK4", **filterkw):
        """members: iterable of dicts with n (grid size, nx = ny), kappa, optional eta, kappah,
        nkappah, L, and optional ``flow`` = a function (x, y, t) -> (u, v) for a time-varying
        flow (evaluated at the member's clock.t before every step, S:708-723); members without
        one take a steady flow from set_flow.  stepper: RK4, ETDRK4, FilteredRK4 or
        FilteredETDRK4 (filter keywords order, innerK, outerK, tol as FourierFlows')."""
        members = list(members)
        st = stepper.upper()
        if st not in _lib.STEPPERS:
            raise ValueError(f"unknown stepper {stepper!r}")
        unknown = set(filterkw) - {"order", "innerK", "outerK", "tol"}
        if unknown or (filterkw and not st.startswith("FILTERED")):
            raise TypeError(f"unexpected keywords {sorted(filterkw)}")
        self._tv = {}
        self._L = []
        arr = (_lib.ptf_desc * len(members))()
        self.sizes = []
        for i, m in enumerate(members):
            d = arr[i]
            n = int(m["n"])
            L = float(m.get("L", 2 * np.pi))
            d.ndim = 2
            d.precision = _lib.PTF_MIXED if precision == "mixed" else _lib.PTF_FP64
            d.n[0], d.n[1], d.n[2] = n, n, 1
            d.L[0], d.L[1], d.L[2] = L, L, 1.0
            d.kappa = float(m["kappa"])
            d.eta = float(m.get("eta", m["kappa"]))
            d.iota = d.kappa
            d.kappah = float(m.get("kappah", 0.0))
            d.nkappah = int(m.get("nkappah", 0))
            d.stepper = _lib.STEPPERS[st]
            d.dt = float(dt)
            if st.startswith("FILTERED"):
                fk = {**dict(order=4, innerK=0.65, outerK=1.0, tol=1e-15), **filterkw}
                d.filter_order = int(fk["order"])
                d.filter_innerK, d.filter_outerK, d.filter_tol = (float(fk["innerK"]),
                                                                  float(fk["outerK"]), float(fk["tol"]))
            if callable(m.get("flow")):
                d.flow_kind = _lib.PTF_FLOW_PER_STEP
                self._tv[i] = m["flow"]
            else:
                d.flow_kind = _lib.PTF_FLOW_STEADY
            self.sizes.append(n)
            self._L.append(L)
        self.precision = "mixed" if precision == "mixed" else "fp64"
        self.stepper = stepper.upper()
        h = C.c_void_p()
        call("ptf_ensemble_create", arr, len(members), C.byref(h))
        self._h = h.value
        self.count = len(members)
        self.steps = 0

    def set_flow(self, i, u, v):
        u = np.ascontiguousarray(u, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        call("ptf_ensemble_set_flow", self._h, i, dptr(u), dptr(v))

    def set_c(self, i, c):
        c = np.ascontiguousarray(c, dtype=np.float64)
        call("ptf_ensemble_set_c", self._h, i, dptr(c))

    def get_c(self, i):
        n = self.sizes[i]
        out = np.empty((n, n))
        call("ptf_ensemble_get_c", self._h, i, dptr(out))
        return out

    def get_sol(self, i):
        n = self.sizes[i]
        out = np.empty((n, n // 2 + 1), np.complex128)
        call("ptf_ensemble_get_sol", self._h, i, dptr(out.view(np.float64)))
        return out

    def step(self, nsteps=1, graph=True):
        """Advance every member nsteps steps.  Members with a time-varying flow get it
        evaluated at their clock.t and uploaded before each step (then the ensemble steps
        one step at a time)."""
        if not self._tv:
            call("ptf_ensemble_step", self._h, int(nsteps), 1 if graph else 0)
        else:
            for _ in range(int(nsteps)):
                for i, f in self._tv.items():
                    t, _ = self.clock(i)
                    n = self.sizes[i]
                    X, Y = _gridpoints2(n, self._L[i])
                    u, v = f(X, Y, t)
                    self.set_flow(i, np.broadcast_to(u, (n, n)), np.broadcast_to(v, (n, n)))
                call("ptf_ensemble_step", self._h, 1, 1 if graph else 0)
        self.steps += int(nsteps)

    def clock(self, i):
        """(t, step) of member i: t advances by the member's dt after every step."""
        t, s = C.c_double(), C.c_int64()
        call("ptf_ensemble_get_clock", self._h, i, C.byref(t), C.byref(s))
        return t.value, s.value

    def set_clock(self, t=0.0, step=0):
        call("ptf_ensemble_set_clock", self._h, float(t), int(step))

    def diagnostics(self, i):
        """FP64 (mass, variance) of member i."""
        m, v = C.c_double(), C.c_double()
        call("ptf_ensemble_diagnostics", self._h, i, C.byref(m), C.byref(v))
        return m.value, v.value

    def set_stream(self, hip_stream):
        call("ptf_ensemble_set_stream", self._h, hip_stream)

    def autotune(self, cache=None):
        """Fastest compiled variant of every launch slot; with ``cache`` (a JSON path) the
        choice is stored there and reused by later runs (see passivetracerflows.tuning)."""
        if cache is None:
            call("ptf_ensemble_autotune", self._h)
            return "tuned"
        from . import tuning
        return tuning.apply_or_tune(self, self.tuning_key(), cache,
                                    tune=lambda: call("ptf_ensemble_autotune", self._h))

    def tuning_key(self):
        from collections import Counter
        sizes = ",".join(f"{c}x{n}^2" for n, c in sorted(Counter(self.sizes).items()))
        return f"ensemble {sizes} {self.precision} {self.stepper}"

    def set_variant(self, i, R, W, occ3=False, par=False, wg=False):
        call("ptf_ensemble_set_variant", self._h, i, R, W, variant_flags(occ3, par, wg))

    def get_variant(self, i):
        """(R, W, occ3, par, wg) of launch slot i."""
        R, W, f = C.c_int32(), C.c_int32(), C.c_int32()
        call("ptf_ensemble_get_variant", self._h, i, C.byref(R), C.byref(W), C.byref(f))
        return R.value, W.value, bool(f.value & 1), bool(f.value & 2), bool(f.value & 4)

    def slots(self):
        n = C.c_int32()
        call("ptf_ensemble_num_slots", self._h, C.byref(n))
        out = []
        for i in range(n.value):
            buf = C.create_string_buffer(256)
            call("ptf_ensemble_slot_name", self._h, i, buf, 256)
            out.append(buf.value.decode())
        return out

    def time_slot(self, i, reps=20):
        ms = C.c_double()
        call("ptf_ensemble_time_slot", self._h, i, reps, C.byref(ms))
        return ms.value

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                _lib.load().ptf_ensemble_destroy(self._h)
                self._h = None
        except Exception:
            pass
