"""ctypes binding of libptf_hip (the C ABI declared in include/ptf.h).

The library is loaded from ``passivetracerflows.jl_amd/lib/libptf_hip.so`` (or
``$PTF_LIB``).  There is no CPU fallback: if the library is missing every
entry point raises ``LibraryNotBuilt``.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PTF_LIB", os.path.join(os.path.dirname(_HERE), "lib", "libptf_hip.so"))

PTF_OK, PTF_ERR_INVALID, PTF_ERR_DEVICE, PTF_ERR_NOMEM, PTF_ERR_UNSUPPORTED = 0, 1, 2, 3, 4
PTF_FP64, PTF_MIXED = 0, 1
PTF_RK4, PTF_ETDRK4, PTF_FILTERED_RK4, PTF_FILTERED_ETDRK4 = 0, 1, 2, 3
STEPPERS = {"RK4": PTF_RK4, "ETDRK4": PTF_ETDRK4, "FILTEREDRK4": PTF_FILTERED_RK4,
            "FILTEREDETDRK4": PTF_FILTERED_ETDRK4}
PTF_FLOW_STEADY, PTF_FLOW_PER_STEP, PTF_FLOW_SPECTRAL, PTF_FLOW_ANALYTIC = 0, 1, 2, 3


class PTFError(RuntimeError):
    """A libptf_hip call returned a non-zero status."""

    def __init__(self, code, msg):
        super().__init__(f"libptf_hip error {code}: {msg}")
        self.code = code


class LibraryNotBuilt(ImportError):
    pass


class ptf_flow_term(C.Structure):
    _fields_ = [
        ("component", C.c_int32), ("fx", C.c_int32), ("fy", C.c_int32), ("fz", C.c_int32),
        ("kx", C.c_double), ("ky", C.c_double), ("kz", C.c_double),
        ("px", C.c_double), ("py", C.c_double), ("pz", C.c_double),
        ("amp", C.c_double), ("time_fn", C.c_int32), ("reserved", C.c_int32),
        ("omega", C.c_double), ("tphase", C.c_double),
    ]


class ptf_desc(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32),
        ("precision", C.c_int32),
        ("n", C.c_int64 * 3),
        ("L", C.c_double * 3),
        ("kappa", C.c_double),
        ("eta", C.c_double),
        ("iota", C.c_double),
        ("kappah", C.c_double),
        ("nkappah", C.c_int32),
        ("stepper", C.c_int32),
        ("dt", C.c_double),
        ("flow_kind", C.c_int32),
        ("filter_order", C.c_int32),
        ("filter_innerK", C.c_double),
        ("filter_outerK", C.c_double),
        ("filter_tol", C.c_double),
    ]


_VP = C.c_void_p
_DP = C.POINTER(C.c_double)
_I32 = C.c_int32
_I64 = C.c_int64

# name -> argtypes (every function returns int status except ptf_last_error)
SIGNATURES = {
    "ptf_version": [C.POINTER(_I32), C.POINTER(_I32)],
    "ptf_device_init": [_I32],
    "ptf_device_synchronize": [],
    "ptf_buffer_alloc": [C.c_uint64, C.POINTER(_VP)],
    "ptf_buffer_free": [_VP],
    "ptf_buffer_upload": [_VP, _VP, C.c_uint64],
    "ptf_buffer_download": [_VP, _VP, C.c_uint64],
    "ptf_plan_create": [_I32, C.POINTER(_I64), _I32, C.POINTER(_VP)],
    "ptf_plan_create_batched": [_I32, C.POINTER(_I64), _I64, _I32, C.POINTER(_VP)],
    "ptf_plan_destroy": [_VP],
    "ptf_plan_set_stream": [_VP, _VP],
    "ptf_rfft": [_VP, _VP, _VP],
    "ptf_irfft": [_VP, _VP, _VP],
    "ptf_irfft_noalloc": [_VP, _VP, _VP],
    "ptf_problem_create": [C.POINTER(ptf_desc), C.POINTER(_VP)],
    "ptf_problem_destroy": [_VP],
    "ptf_problem_set_stream": [_VP, _VP],
    "ptf_problem_get_stream": [_VP, C.POINTER(_VP)],
    "ptf_set_flow": [_VP, _DP, _DP, _DP],
    "ptf_set_spectral_flow": [_VP, _DP, _DP],
    "ptf_set_analytic_flow": [_VP, C.POINTER(ptf_flow_term), _I32],
    "ptf_set_c": [_VP, _DP],
    "ptf_step": [_VP, _I64],
    "ptf_problem_set_graph": [_VP, _I32],
    "ptf_get_clock": [_VP, _DP, C.POINTER(_I64)],
    "ptf_set_clock": [_VP, C.c_double, _I64],
    "ptf_get_sol": [_VP, _DP],
    "ptf_set_sol": [_VP, _DP],
    "ptf_updatevars": [_VP, _DP],
    "ptf_calcN": [_VP, _DP, _DP],
    "ptf_get_L": [_VP, _DP],
    "ptf_diagnostics": [_VP, _DP, _DP],
    "ptf_problem_sol_ptr": [_VP, C.POINTER(_VP)],
    "ptf_autotune": [_VP],
    "ptf_problem_num_slots": [_VP, C.POINTER(_I32)],
    "ptf_problem_slot_name": [_VP, _I32, C.c_char_p, _I32],
    "ptf_problem_time_slot": [_VP, _I32, _I32, _DP],
    "ptf_step_profiled": [_VP, _I64, _DP],
    "ptf_dealias": [_VP, _VP, C.c_double],
    "ptf_step_record": [_VP, _I64, _I64, _I64, C.POINTER(_I64), _DP, C.POINTER(_I64), _DP, _DP,
                        _DP],
    "ptf_problem_set_variant": [_VP, _I32, _I32, _I32, _I32],
    "ptf_problem_get_variant": [_VP, _I32, C.POINTER(_I32), C.POINTER(_I32), C.POINTER(_I32)],
    "ptf_ensemble_create": [C.POINTER(ptf_desc), _I32, C.POINTER(_VP)],
    "ptf_ensemble_destroy": [_VP],
    "ptf_ensemble_set_stream": [_VP, _VP],
    "ptf_ensemble_set_flow": [_VP, _I32, _DP, _DP],
    "ptf_ensemble_set_c": [_VP, _I32, _DP],
    "ptf_ensemble_get_c": [_VP, _I32, _DP],
    "ptf_ensemble_get_sol": [_VP, _I32, _DP],
    "ptf_ensemble_step": [_VP, _I64, _I32],
    "ptf_ensemble_autotune": [_VP],
    "ptf_ensemble_num_slots": [_VP, C.POINTER(_I32)],
    "ptf_ensemble_slot_name": [_VP, _I32, C.c_char_p, _I32],
    "ptf_ensemble_time_slot": [_VP, _I32, _I32, _DP],
    "ptf_ensemble_set_variant": [_VP, _I32, _I32, _I32, _I32],
    "ptf_ensemble_get_clock": [_VP, _I32, _DP, C.POINTER(_I64)],
    "ptf_ensemble_set_clock": [_VP, C.c_double, _I64],
    "ptf_ensemble_diagnostics": [_VP, _I32, _DP, _DP],
    "ptf_ensemble_get_variant": [_VP, _I32, C.POINTER(_I32), C.POINTER(_I32), C.POINTER(_I32)],
}

_lib = None


def load(path: str | None = None):
    """Load (once) and return the ctypes library with prototypes set."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise LibraryNotBuilt(
            f"libptf_hip is not built ({p} missing); run passivetracerflows.jl_amd/build.py")
    lib = C.CDLL(p)
    lib.ptf_last_error.restype = C.c_char_p
    lib.ptf_last_error.argtypes = []
    for name, args in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = C.c_int
        f.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != PTF_OK:
        msg = load().ptf_last_error()
        raise PTFError(rc, msg.decode() if msg else "")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))


def dptr(a):
    """ctypes double* of a C-contiguous float64 numpy array."""
    return a.ctypes.data_as(_DP)
