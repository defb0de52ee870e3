"""ptf_irfft_noalloc (the inverse transform that allocates nothing) against ptf_irfft.

Both run the same kernels; the allocation-free entry point only takes its intermediate
spectrum from the library's static device array instead of the plan's scratch buffer,
so the outputs must be bitwise equal.  Two calls in a row with different plans, the
second smaller than the first (the static array keeps the first call's contents), in
FP64 and in the mixed precision; a plan whose spectra exceed the static array is
refused before anything is launched.
"""
import ctypes as C

import numpy as np
import pytest

from passivetracerflows import _lib

pytestmark = pytest.mark.gpu


def _check(lib, rc):
    assert rc == 0, lib.ptf_last_error().decode()


def _plan(lib, ndim, n, prec):
    h = C.c_void_p()
    _check(lib, lib.ptf_plan_create(ndim, (C.c_int64 * 3)(*n), prec, C.byref(h)))
    return h


def _dev(lib, nbytes):
    p = C.c_void_p()
    _check(lib, lib.ptf_buffer_alloc(C.c_uint64(nbytes), C.byref(p)))
    return p


def _case(lib, ndim, n, prec, seed):
    nx, ny, nz = n
    rt = np.float64 if prec == _lib.PTF_FP64 else np.float32
    ct = np.complex128 if prec == _lib.PTF_FP64 else np.complex64
    shape = (nz, ny, nx)[3 - ndim:]
    spec_shape = shape[:-1] + (nx // 2 + 1,)
    rng = np.random.default_rng(seed)
    spec = (rng.standard_normal(spec_shape) + 1j * rng.standard_normal(spec_shape)).astype(ct)
    print(f"irfft_noalloc ndim={ndim} n={n} prec={prec}", flush=True)
    plan = _plan(lib, ndim, n, prec)
    din = _dev(lib, spec.nbytes)
    _check(lib, lib.ptf_buffer_upload(din, spec.ctypes.data, spec.nbytes))
    outs = []
    for fn in (lib.ptf_irfft, lib.ptf_irfft_noalloc):
        out = np.zeros(shape, rt)
        dout = _dev(lib, out.nbytes)
        _check(lib, fn(plan, din, dout))
        _check(lib, lib.ptf_device_synchronize())
        _check(lib, lib.ptf_buffer_download(out.ctypes.data, dout, out.nbytes))
        _check(lib, lib.ptf_buffer_free(dout))
        outs.append(out)
    back = np.zeros(spec_shape, ct)
    _check(lib, lib.ptf_buffer_download(back.ctypes.data, din, back.nbytes))
    _check(lib, lib.ptf_buffer_free(din))
    _check(lib, lib.ptf_plan_destroy(plan))
    assert np.array_equal(back, spec)               # the input is preserved
    assert np.all(np.isfinite(outs[0])) and np.abs(outs[0]).max() > 0
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("prec", [_lib.PTF_FP64, _lib.PTF_MIXED])
def test_noalloc_equals_irfft_two_calls(dev, prec):
    lib = _lib.load()
    _case(lib, 3, (64, 128, 64), prec, seed=1)   # first call: 64^2 x 128 spectra
    _case(lib, 2, (32, 16, 1), prec, seed=2)     # second call: a smaller plan
    _case(lib, 1, (64, 1, 1), prec, seed=3)      # 1D: no intermediate spectrum at all


def test_noalloc_rejects_oversized_plan(dev):
    lib = _lib.load()
    plan = _plan(lib, 3, (256, 256, 256), _lib.PTF_FP64)   # 135 MB of half spectrum
    try:
        rc = lib.ptf_irfft_noalloc(plan, C.c_void_p(16), C.c_void_p(16))
        assert rc == _lib.PTF_ERR_INVALID
        assert "static scratch" in lib.ptf_last_error().decode()
    finally:
        _check(lib, lib.ptf_plan_destroy(plan))
