"""Host-side guards that run before anything reaches the GPU (CPU tests).

* The HIP-runtime check: a process that imports a framework bundling its own ROCm
  release (torch here) before libptf_hip binds the library to that runtime; every
  entry point that first touches the device must refuse it with PTF_ERR_DEVICE.
* The plan interface (mul!/ldiv!/dealias!) checks the shape, precision and owner of
  its arrays before it launches anything.
"""
import os
import subprocess
import sys
import textwrap

import numpy as np
import pytest

import passivetracerflows as ptf
from passivetracerflows import _lib
from passivetracerflows import fourierflows as FF

PKG = os.path.dirname(os.path.dirname(os.path.abspath(_lib.__file__)))


def _lib_built():
    return os.path.exists(_lib.LIB_PATH)


def _probe(import_torch: bool):
    code = textwrap.dedent(f"""
        import ctypes, sys
        if {import_torch!r}:
            import torch  # noqa: F401  (loads torch's bundled libamdhip64 first)
        lib = ctypes.CDLL({_lib.LIB_PATH!r})
        lib.ptf_last_error.restype = ctypes.c_char_p
        rc = lib.ptf_device_init(0)
        print(rc)
        print(lib.ptf_last_error().decode())
        p = ctypes.c_void_p()
        rc2 = lib.ptf_buffer_alloc(ctypes.c_uint64(64), ctypes.byref(p))
        print(rc2)
        print(lib.ptf_last_error().decode())
    """)
    env = dict(os.environ, HIP_VISIBLE_DEVICES=os.environ.get("HIP_VISIBLE_DEVICES", "0"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300,
                       env=env)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    return int(lines[0]), lines[1], int(lines[2]), lines[3]


@pytest.mark.skipif(not _lib_built(), reason="libptf_hip not built")
def test_runtime_guard_fires_after_torch():
    pytest.importorskip("torch")
    rc, msg, rc2, msg2 = _probe(import_torch=True)
    assert rc == _lib.PTF_ERR_DEVICE, msg
    assert "HIP runtime mismatch" in msg and "torch" in msg, msg
    assert rc2 == _lib.PTF_ERR_DEVICE and "HIP runtime mismatch" in msg2, msg2


@pytest.mark.skipif(not _lib_built(), reason="libptf_hip not built")
def test_runtime_guard_silent_without_torch():
    rc, msg, _, _ = _probe(import_torch=False)
    assert "HIP runtime mismatch" not in msg, msg


class _FakeArray(FF.DeviceArray):
    """A DeviceArray that owns no device memory (the guard must fire before any use)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = 0x1000

    def __del__(self):
        pass


def _fake_plan(ndim, n, T, nlayers=1):
    dev = object.__new__(FF.ROCm)
    dev.ordinal = 0
    g = {1: FF.OneDGrid, 2: FF.TwoDGrid, 3: FF.ThreeDGrid}[ndim](dev, nx=n, T=T)
    plan = object.__new__(FF.RFFTPlan)
    plan.grid = g
    plan.nlayers = nlayers
    plan._h = None
    return g, plan


def _never_called(*a, **k):
    raise AssertionError("a library call was made before the arrays were checked")


@pytest.mark.parametrize("ndim,T", [(1, np.float64), (2, np.float64), (3, np.float32)])
def test_plan_arrays_are_checked_before_launch(monkeypatch, ndim, T):
    monkeypatch.setattr(FF, "call", _never_called)
    g, plan = _fake_plan(ndim, 16, T)
    ct = np.complex128 if T is np.float64 else np.complex64
    other_t = np.float32 if T is np.float64 else np.float64
    good_r, good_c = _FakeArray(g.shape, T), _FakeArray(g.spec_shape, ct)
    small = _FakeArray((4,) * ndim, T)
    bad_cases = [
        (good_c, small),                               # too small an input
        (good_c, _FakeArray(g.shape, other_t)),        # wrong precision
        (_FakeArray(g.shape, T), good_r),              # real array where complex is due
        (good_c, np.zeros(g.shape, T)),                # a host array
    ]
    for out, a in bad_cases:
        with pytest.raises((ValueError, TypeError)):
            ptf.mul_(out, plan, a)
    with pytest.raises((ValueError, TypeError)):
        ptf.ldiv_(good_r, plan, _FakeArray(g.shape, ct))
    with pytest.raises((ValueError, TypeError)):
        ptf.ldiv_(_FakeArray(g.spec_shape, T), plan, good_c)
    with pytest.raises((ValueError, TypeError)):
        ptf.dealias_(_FakeArray(g.shape, ct), g, plan)
    # a layered plan wants the leading layer axis
    _, lplan = _fake_plan(ndim, 16, T, nlayers=3)
    with pytest.raises(ValueError):
        ptf.mul_(good_c, lplan, good_r)


class _ForeignTensor:
    """Stands for a framework's device tensor (an object exposing data_ptr())."""

    shape = (16, 16)
    dtype = np.float64

    def data_ptr(self):
        return 0x2000


def test_foreign_tensor_rejected(monkeypatch):
    monkeypatch.setattr(FF, "call", _never_called)
    g, plan = _fake_plan(2, 16, np.float64)
    with pytest.raises(TypeError):
        ptf.mul_(_FakeArray(g.spec_shape, np.complex128), plan, _ForeignTensor())
