"""Build libptf_hip.so for gfx950 in-tree (passivetracerflows.jl_amd/lib/).

    python passivetracerflows.jl_amd/build.py [--jobs N] [--incremental]

The kernel registry (csrc/ptf_registry.hip) is compiled once per
(precision, kernel kind) so the units build in parallel; csrc/ptf_member.hip holds
the whole-member step kernels; csrc/ptf_api.hip is the host runtime.  All HIP code is built with packed-FP32 instructions off.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "_build")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIBNAME = "libptf_hip.so"

ARCH = os.environ.get("PTF_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", CSRC, "-I", INCLUDE, "-Wno-unused-command-line-argument",
]
KINDS = [0, 1, 2, 3, 4, 5, 6, 7, 8, 10, 11, 12, 13, 14]
# parts per kind: the line lengths of the heavy kinds (stage epilogues, advection) are
# dealt round-robin over several objects so that they compile in parallel (at most 6)
PARTS = {3: 2, 4: 4, 5: 2, 6: 6, 7: 4, 8: 6, 12: 2, 13: 3, 14: 3}
PRECS = [("double", "f64"), ("float", "f32")]
HEADERS = ["ptf_device.hpp", "ptf_kernels.hpp"]
API_HEADERS = HEADERS + ["ptf_setup.hpp"]


def _units():
    units = []
    for t, tn in PRECS:
        for k in KINDS:
            nparts = PARTS.get(k, 1)
            for part in range(nparts):
                obj = os.path.join(OBJDIR, f"reg_{tn}_{k}_{part}.o")
                cmd = [HIPCC, *COMMON, f"-DPTF_T={t}", f"-DPTF_TNAME={tn}", f"-DPTF_KIND={k}",
                       f"-DPTF_PART={part}", f"-DPTF_NPARTS={nparts}",
                       "-c", os.path.join(CSRC, "ptf_registry.hip"), "-o", obj]
                units.append((obj, cmd, ["ptf_registry.hip", *HEADERS]))
    # the whole-member step kernels (small 2D ensemble members)
    obj = os.path.join(OBJDIR, "ptf_member.o")
    units.append((obj, [HIPCC, *COMMON, "-c", os.path.join(CSRC, "ptf_member.hip"), "-o", obj],
                  ["ptf_member.hip", "ptf_member.hpp", *HEADERS]))
    # the heaviest units first, so that they do not end up on the critical path last
    units.sort(key=lambda u: -_weight(u[0]))
    obj = os.path.join(OBJDIR, "ptf_api.o")
    units.append((obj, [HIPCC, *COMMON, "-c", os.path.join(CSRC, "ptf_api.hip"), "-o", obj],
                  ["ptf_api.hip", *API_HEADERS, "ptf_member.hpp", os.path.join(INCLUDE, "ptf.h")]))
    return units


def _weight(obj):
    name = os.path.basename(obj)
    if not name.startswith("reg_"):
        return 0
    k = int(name.split("_")[2])
    return {6: 10, 8: 10, 4: 6, 7: 6, 12: 4, 13: 4, 14: 4, 3: 2, 5: 2}.get(k, 1) / PARTS.get(k, 1)


def _stale(obj, deps):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(os.path.join(CSRC, d)) > t for d in deps)


def _run(cmd):
    t0 = time.time()
    p = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in (p.stderr or "").splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    return p.returncode, err, time.time() - t0


def build(jobs: int | None = None, incremental: bool = False, verbose: bool = True) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    units = _units()
    todo = [(o, c) for o, c, d in units if not incremental or _stale(o, d)]
    t0 = time.time()
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {ex.submit(_run, c): o for o, c in todo}
        for f in cf.as_completed(futs):
            rc, err, dt = f.result()
            if rc != 0:
                raise RuntimeError(f"hipcc failed for {futs[f]}:\n{err}")
            if verbose:
                print(f"  built {os.path.basename(futs[f])} in {dt:.1f}s", flush=True)
    lib = os.path.join(LIBDIR, LIBNAME)
    tmp = lib + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp,
           *[o for o, _, _ in units]]
    rc, err, _ = _run(cmd)
    if rc != 0:
        raise RuntimeError(f"link failed:\n{err}")
    os.replace(tmp, lib)
    if verbose:
        print(f"libptf_hip: {lib} ({time.time() - t0:.1f}s)", flush=True)
    return lib


def build_asan(verbose: bool = True) -> str:
    """Host-side AddressSanitizer build of the C-ABI runtime: csrc/ptf_api.hip compiled with
    ASan on its host code only (the device code and the kernel registry objects are the
    regular ones), linked into lib/libptf_hip_asan.so.  Needs a regular build's objects.
    Load it with the clang ASan runtime preloaded (see tests/test_sanitizer.py); it is for
    the CPU tests of the boundary (argument checks, error paths, host buffers)."""
    units = _units()
    regs = [o for o, _, _ in units if os.path.basename(o).startswith(("reg_", "ptf_member"))]
    missing = [o for o in regs if not os.path.exists(o)]
    if missing:
        raise RuntimeError("build the library first (missing " + os.path.basename(missing[0]) + ")")
    obj = os.path.join(OBJDIR, "ptf_api_asan.o")
    flags = [f for f in COMMON if f != "-O3"]
    cmd = [HIPCC, *flags, "-O1", "-g", "-Xarch_host", "-fsanitize=address",
           "-Xarch_host", "-fno-omit-frame-pointer", "-c", os.path.join(CSRC, "ptf_api.hip"), "-o", obj]
    rc, err, dt = _run(cmd)
    if rc != 0:
        raise RuntimeError(f"hipcc (asan) failed:\n{err}")
    lib = os.path.join(LIBDIR, "libptf_hip_asan.so")
    clangxx = os.path.join(os.path.dirname(os.path.realpath(HIPCC)), "..", "lib", "llvm", "bin", "clang++")
    if not os.path.exists(clangxx):
        clangxx = "/opt/rocm/lib/llvm/bin/clang++"
    link = [clangxx, "-shared", "-fPIC", "-fsanitize=address", "-shared-libasan", "-o", lib, obj, *regs,
            "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]
    rc, err, _ = _run(link)
    if rc != 0:
        raise RuntimeError(f"asan link failed:\n{err}")
    if verbose:
        print(f"libptf_hip_asan: {lib} ({dt:.1f}s)", flush=True)
    return lib


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("--asan", action="store_true",
                    help="also link lib/libptf_hip_asan.so (host AddressSanitizer build of the runtime)")
    a = ap.parse_args()
    try:
        build(a.jobs, a.incremental)
        if a.asan:
            build_asan()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
