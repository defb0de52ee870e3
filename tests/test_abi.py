"""CPU tests of the C ABI boundary and the host-side logic (no GPU compute).

* the library loads and exports every entry point include/ptf.h declares;
* descriptor validation and NULL-handle errors are returned before any device
  call, with a message in ptf_last_error();
* the Python mirror's argument handling and the workload generators.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from passivetracerflows import _lib
from passivetracerflows import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ptf.h")


def header_functions():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(ptf_\w+)\s*\(", src, re.M)))


def test_header_matches_binding():
    names = header_functions()
    assert "ptf_step" in names and "ptf_rfft" in names
    bound = set(_lib.SIGNATURES) | {"ptf_last_error"}
    assert set(names) == bound, set(names) ^ bound


def test_library_exports_every_symbol():
    lib = _lib.load()
    for n in header_functions():
        assert hasattr(lib, n), n


def test_desc_layout_matches_header(tmp_path):
    """Field offsets of the ctypes ptf_desc equal the C compiler's for include/ptf.h."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    fields = [f[0] for f in _lib.ptf_desc._fields_]
    src = tmp_path / "off.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "ptf.h"\nint main(void){\n'
                   + "".join(f'printf("%zu\\n", offsetof(ptf_desc, {f}));\n' for f in fields)
                   + 'printf("%zu\\n", sizeof(ptf_desc));return 0;}\n')
    exe = tmp_path / "off"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    assert out[:-1] == [getattr(_lib.ptf_desc, f).offset for f in fields]
    assert out[-1] == C.sizeof(_lib.ptf_desc)


def _desc(**kw):
    d = _lib.ptf_desc()
    d.ndim = kw.get("ndim", 2)
    d.precision = kw.get("precision", _lib.PTF_FP64)
    n = kw.get("n", (64, 64, 1))
    for i in range(3):
        d.n[i] = n[i]
        d.L[i] = kw.get("L", 2 * np.pi)
    d.kappa = d.eta = d.iota = 0.01
    d.kappah = 0.0
    d.nkappah = kw.get("nkappah", 0)
    d.stepper = kw.get("stepper", _lib.PTF_RK4)
    d.dt = kw.get("dt", 0.01)
    d.flow_kind = kw.get("flow_kind", _lib.PTF_FLOW_STEADY)
    return d


@pytest.mark.parametrize("kw,code,msg", [
    (dict(ndim=4), _lib.PTF_ERR_INVALID, "ndim"),
    (dict(n=(63, 64, 1)), _lib.PTF_ERR_INVALID, "even"),
    (dict(n=(98, 64, 1)), _lib.PTF_ERR_UNSUPPORTED, "no compiled transform"),
    (dict(n=(64, 2002, 1)), _lib.PTF_ERR_UNSUPPORTED, "no compiled transform"),   # 2 * 7 * 11 * 13
    (dict(n=(8, 64, 1)), _lib.PTF_ERR_UNSUPPORTED, "nx must be"),
    (dict(n=(64, 4096, 1)), _lib.PTF_ERR_UNSUPPORTED, "ny and nz"),
    (dict(dt=0.0), _lib.PTF_ERR_INVALID, "dt"),
    (dict(dt=float("nan")), _lib.PTF_ERR_INVALID, "dt"),
    (dict(precision=7), _lib.PTF_ERR_INVALID, "precision"),
    (dict(stepper=5), _lib.PTF_ERR_INVALID, "stepper"),
    (dict(nkappah=-1), _lib.PTF_ERR_INVALID, "nkappah"),
    (dict(ndim=3, n=(16, 16, 16), flow_kind=_lib.PTF_FLOW_SPECTRAL), _lib.PTF_ERR_UNSUPPORTED,
     "2D only"),
    (dict(L=-1.0), _lib.PTF_ERR_INVALID, "lengths"),
])
def test_problem_create_rejects(kw, code, msg):
    lib = _lib.load()
    d = _desc(**kw)
    h = C.c_void_p()
    rc = lib.ptf_problem_create(C.byref(d), C.byref(h))
    assert rc == code
    assert h.value is None
    assert msg in lib.ptf_last_error().decode()


@pytest.mark.parametrize("n", [(96, 192, 1), (384, 96, 1), (640, 1000, 1), (1000, 64, 1)])
def test_problem_create_accepts_mixed_radix_sizes(n):
    """2^a 3^b 5^c sizes pass validation (on a machine without a GPU the creation then fails
    at the first device call, never with PTF_ERR_UNSUPPORTED or PTF_ERR_INVALID)."""
    lib = _lib.load()
    for prec in (_lib.PTF_FP64, _lib.PTF_MIXED):
        d = _desc(n=n, precision=prec)
        h = C.c_void_p()
        rc = lib.ptf_problem_create(C.byref(d), C.byref(h))
        msg = lib.ptf_last_error().decode()
        assert rc not in (_lib.PTF_ERR_UNSUPPORTED, _lib.PTF_ERR_INVALID), msg
        if rc == _lib.PTF_OK:
            lib.ptf_problem_destroy(h)


def test_null_handles_rejected():
    lib = _lib.load()
    assert lib.ptf_step(None, 1) == _lib.PTF_ERR_INVALID
    assert lib.ptf_set_c(None, None) == _lib.PTF_ERR_INVALID
    assert lib.ptf_rfft(None, None, None) == _lib.PTF_ERR_INVALID
    assert lib.ptf_problem_create(None, None) == _lib.PTF_ERR_INVALID
    assert lib.ptf_ensemble_create(None, 0, None) == _lib.PTF_ERR_INVALID
    assert lib.ptf_ensemble_step(None, 1, 1) == _lib.PTF_ERR_INVALID
    h = C.c_void_p()
    assert lib.ptf_plan_create(2, None, 0, C.byref(h)) == _lib.PTF_ERR_INVALID
    n = (C.c_int64 * 3)(42, 42, 1)      # 42 = 2 * 3 * 7: no compiled transform
    assert lib.ptf_plan_create(2, n, 0, C.byref(h)) == _lib.PTF_ERR_UNSUPPORTED


def test_python_error_raises():
    with pytest.raises(_lib.PTFError) as e:
        _lib.call("ptf_step", None, 1)
    assert e.value.code == _lib.PTF_ERR_INVALID


def test_workload_shapes_and_counts():
    counts = W.ensemble_counts(512)
    assert counts == [273, 137, 68, 34]
    assert sum(n * n * c for n, c in zip(W.ENSEMBLE_SIZES, counts)) == 16_732_160
    c = W.cellular_2d(32, nsteps=1)
    u, v = c["flow"]
    # divergence free: d_x u + d_y v = 0 spectrally
    k = np.fft.fftfreq(32, 1 / 32)
    uh, vh = np.fft.fft2(u), np.fft.fft2(v)
    div = 1j * k[None, :] * uh + 1j * k[:, None] * vh
    assert np.abs(div).max() < 1e-10 * np.abs(uh).max()
    f = W.random_field_3d(16)
    assert abs(f.mean()) < 1e-12 and abs(f.std() - 1) < 1e-12
    psih, om = W.turbulent_psih(64)
    assert psih.shape == (64, 33) and om.shape == (64, 33)
    uu, vv = W.spectral_flow_at(psih, om, 0.0, 64)
    assert abs(np.sqrt(np.mean(uu ** 2 + vv ** 2)) - 1) < 1e-12


def test_mirror_keywords_without_device():
    from passivetracerflows import traceradvectiondiffusion as TAD
    assert TAD.noflow(np.pi) == 0
    p = TAD.ConstDiffTimeVaryingFlowParams(0.1, 0.2, lambda x, y, t: 0, lambda x, y, t: 0)
    assert (p.κ, p.η, p.κh, p.nκh) == (0.1, 0.2, 0.0, 0)
    assert TAD._flow_kind(p) == _lib.PTF_FLOW_PER_STEP
    sp = TAD.ConstDiffSpectralFlowParams(0.0, 0.0, 1e-12, 4, TAD.SpectralStreamfunctionFlow(
        np.zeros((8, 5)), np.zeros((8, 5))))
    assert TAD._flow_kind(sp) == _lib.PTF_FLOW_SPECTRAL
