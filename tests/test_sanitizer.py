"""The C-ABI runtime's host code under AddressSanitizer (CPU test).

passivetracerflows.jl_amd/build.py --asan links lib/libptf_hip_asan.so: csrc/ptf_api.hip
with ASan on its host code (device code unchanged).  The boundary tests that need no GPU
(argument checks and error paths, descriptor validation, the HIP-runtime guard, the
plan-array checks, the Julia binding's prototypes) run against it in a child process
with the clang ASan runtime preloaded; any heap overflow, use-after-free or double free in
those paths fails the run.  Leak checking is off (the Python interpreter itself leaks at
exit by design).
"""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "passivetracerflows.jl_amd")
sys.path.insert(0, PKG)


def _asan_runtime():
    hits = sorted(glob.glob("/opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return hits[-1] if hits else None


def test_abi_paths_clean_under_asan():
    rt = _asan_runtime()
    if rt is None:
        pytest.skip("clang ASan runtime not found")
    if not os.path.exists(os.path.join(PKG, "_build", "ptf_api.o")):
        pytest.skip("regular build objects missing (run build.py first)")
    import build as B
    lib = B.build_asan(verbose=False)
    env = dict(os.environ, LD_PRELOAD=rt, PTF_LIB=lib,
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0:exitcode=66")
    env.pop("PYTEST_ADDOPTS", None)
    # the instrumented library is the one the tests load
    probe = ("import sys; sys.path.insert(0, %r); from passivetracerflows import _lib; _lib.load(); "
             "print(any('libptf_hip_asan.so' in l for l in open('/proc/self/maps')))" % PKG)
    r0 = subprocess.run([sys.executable, "-c", probe], cwd=ROOT, env=env, capture_output=True,
                        text=True, timeout=300)
    assert r0.stdout.strip().endswith("True"), r0.stdout + r0.stderr
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                        "tests/test_abi.py", "tests/test_host_guards.py",
                        "tests/test_julia_shim.py", "-k", "not runtime_guard_fires"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    tail = (r.stdout[-3000:] + r.stderr[-3000:])
    assert "AddressSanitizer" not in r.stderr, tail
    assert r.returncode == 0, tail
    assert " passed" in r.stdout, tail
