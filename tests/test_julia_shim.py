"""The Julia binding (julia/PTFHip/src/PTFHip.jl) against include/ptf.h (CPU tests).

Julia is not installed in this image, so the binding cannot run here; what can be
checked without it is that it binds the C ABI correctly:

* ``PtfDesc`` and ``PtfFlowTerm`` have the fields of ``ptf_desc`` / ``ptf_flow_term`` in
  the same order with types of the same size, and the C layout Julia gives an
  ``isbits`` struct (natural alignment, as C) puts every field at the offset the C
  compiler reports for the header;
* every ``ccall((:ptf_*, libptf), ...)`` names a function the header declares and the
  library exports, with as many argument types as the C prototype has parameters;
* the methods that make the reference's scripts run on the device dispatch on this
  package's types only (no method of FourierFlows or PassiveTracerFlows is redefined).
"""
import os
import re
import shutil
import subprocess

import pytest

from passivetracerflows import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = os.path.join(ROOT, "passivetracerflows.jl_amd", "julia", "PTFHip", "src", "PTFHip.jl")
HEADER = os.path.join(ROOT, "include", "ptf.h")

JULIA_C = {"Int32": ("int32_t", 4), "Int64": ("int64_t", 8), "Float64": ("double", 8),
           "NTuple{3,Int64}": ("int64_t[3]", 24), "NTuple{3,Float64}": ("double[3]", 24)}
ALIGN = {"Int32": 4, "Int64": 8, "Float64": 8, "NTuple{3,Int64}": 8, "NTuple{3,Float64}": 8}


def julia_struct(name):
    src = open(JL).read()
    m = re.search(rf"^struct {name}\n(.*?)^end", src, re.M | re.S)
    assert m, name
    fields = []
    for line in m.group(1).splitlines():
        line = line.split("#")[0].strip()
        if line:
            f, t = line.split("::")
            fields.append((f.strip(), t.strip()))
    return fields


def c_offsets(struct, fields, tmp_path):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / f"{struct}.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "ptf.h"\nint main(void){\n'
                   + "".join(f'printf("%zu %zu\\n", offsetof({struct}, {f}), sizeof((({struct}*)0)->{f}));\n'
                             for f in fields)
                   + f'printf("%zu 0\\n", sizeof({struct}));return 0;}}\n')
    exe = tmp_path / struct
    subprocess.run([cc, "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    rows = [tuple(map(int, l.split())) for l in
            subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines()]
    return rows[:-1], rows[-1][0]


def julia_layout(fields):
    """Offsets of an isbits Julia struct: each field at the next multiple of its alignment."""
    off, out, amax = 0, [], 1
    for _, t in fields:
        a, size = ALIGN[t], JULIA_C[t][1]
        off = (off + a - 1) // a * a
        out.append((off, size))
        off += size
        amax = max(amax, a)
    return out, (off + amax - 1) // amax * amax


@pytest.mark.parametrize("jl,c", [("PtfDesc", "ptf_desc"), ("PtfFlowTerm", "ptf_flow_term")])
def test_struct_layout_matches_header(jl, c, tmp_path):
    fields = julia_struct(jl)
    for _, t in fields:
        assert t in JULIA_C, t
    c_rows, c_size = c_offsets(c, [f for f, _ in fields], tmp_path)
    j_rows, j_size = julia_layout(fields)
    assert j_rows == c_rows, list(zip([f for f, _ in fields], j_rows, c_rows))
    assert j_size == c_size
    # and the ctypes binding of the Python mirror agrees on the field order
    py = {"ptf_desc": _lib.ptf_desc, "ptf_flow_term": _lib.ptf_flow_term}[c]
    assert [f for f, _ in py._fields_] == [f for f, _ in fields]


def _c_prototypes():
    src = open(HEADER).read()
    protos = {}
    for m in re.finditer(r"^\s*(?:int|const char\*)\s+(ptf_\w+)\s*\(([^)]*)\)\s*;", src, re.M | re.S):
        args = m.group(2).strip()
        protos[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return protos


def test_every_ccall_is_declared_exported_and_arity_matches():
    src = open(JL).read()
    protos = _c_prototypes()
    lib = _lib.load()
    calls = re.findall(r"ccall\(\(:(ptf_\w+),\s*libptf\),\s*\w+,\s*\(", src)
    assert len(calls) >= 20
    for m in re.finditer(r"ccall\(\(:(ptf_\w+),\s*libptf\),\s*\w+,\s*\(", src):
        name = m.group(1)
        assert name in protos, name
        assert hasattr(lib, name), name
        # the argument-type tuple that follows: count top-level commas
        i, depth, n_args, start = m.end(), 1, 0, m.end()
        while depth:
            ch = src[i]
            if ch in "({[":
                depth += 1
            elif ch in ")}]":
                depth -= 1
            elif ch == "," and depth == 1:
                n_args += 1
            i += 1
        body = src[start:i - 1].strip()
        n = 0 if body == "" else (n_args + 1 if not body.endswith(",") else n_args)
        assert n == protos[name], (name, n, protos[name])


def test_methods_dispatch_on_package_types_only():
    src = open(JL).read()
    # every method added to FourierFlows / PassiveTracerFlows / Base names a PTFHip type
    own = ("ROCm", "PtfSol", "PtfArray", "PtfRFFTPlan", "PtfTimeStepper", "FlowParams")
    for m in re.finditer(r"^function ((?:FourierFlows|TAD|Base)\.\w+)\((.*?)\)(?: where .*)?$", src, re.M):
        assert any(t in m.group(2) for t in own), m.group(0)
    for m in re.finditer(r"^((?:FourierFlows|TAD|Base)\.\w+)\((.*?)\)(?: where [^=]*)? =", src, re.M):
        assert any(t in m.group(2) for t in own), m.group(0)
    # the reference entry points a script calls all route here
    for needed in ("FourierFlows.TwoDGrid(dev::ROCm", "FourierFlows.TimeStepper(", "Base.zeros(::ROCm",
                   "FourierFlows.stepforward!(sol::PtfSol", "TAD.set_c!(sol::PtfSol",
                   "TAD.updatevars!(params::FlowParams, vars, grid, sol::PtfSol)"):
        assert needed in src, needed
