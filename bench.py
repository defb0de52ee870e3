"""Benchmark: TracerAdvectionDiffusion time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extra]

Headline workload: 3D steady ABC flow, 256^3, RK4, dt = 5e-3, kappa = 1e-3,
FP64 plain form (inputs resident in HBM before the timed region).  A "step"
is one full RK4 step (4 calcN! evaluations + stage updates) of that problem;
value = grid points x steps / wall seconds.

Also reported (``forms``): the same problem in the mixed FP32 form, the
512-member FP32 ensemble (64^2..512^2, hipGraph replay), the 2D 256^2
cellular flow (FP64, mixed) and the 2D 2048^2 random-phase ETDRK4 flow.
The ``roofline`` object is for the kernel that takes the most time per step
of the headline workload; its bytes come from the pass model of each kernel
(see the KERNEL_BYTES table).  ``cpu_baseline`` times the FP64 oracle
restatement (scipy pocketfft threads + numpy) on a bounded number of steps.
"""
from __future__ import annotations

import argparse
import faulthandler
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "passivetracerflows.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "grid-point-steps/s, FP64 plain vs FP32 batched at relL2≤1e-5; %HBM roofline"
# the autotuner's per-slot choices, stored by the first tuned run and applied by later ones
# (so a profiled run times the same variants as the benchmark); --retune refreshes them
TUNING = os.path.join(ROOT, "passivetracerflows.jl_amd", "passivetracerflows", "tuning_gfx950.json")
HBM_PEAK_GBS = 8000.0
# fixed yardstick for the whole step: every FFT axis pass reads and writes its fields once,
# flows read once per stage, RK4 state 4/6/6/5 S-fields per stage: 77 S + 12 R per step
# (S = 16 B x 129/128 per point, R = 8 B per point at 256^3 FP64)
PASS_MODEL_B_PER_GPS = 716.8


def kernel_bytes(name: str, nS: int, nR: int, cb: int, rb: int, stage: int, stepper: str = "RK4"):
    """Algorithmic HBM bytes of one launch (pass model): every field a kernel must
    read or write once.  cb/rb = bytes per complex/real element (cb = 8: the mixed form,
    whose compensated state adds low parts)."""
    S, R = nS * cb, nR * rb
    C = nS * 8       # an FP64 ETDRK4 coefficient array (S-shaped, real)
    fam = name.split(" [")[0].split("_", 1)[1] if name.startswith("stage") else name.split(" [")[0]
    if fam.startswith("prologue_"):
        return 3 * S
    if fam == "z_inverse_deriv":
        return 3 * S                     # read c, write IFFT_z c, IFFT_z(i m c)
    if fam == "y_inverse_deriv2":
        return 5 * S                     # read 2, write 3
    if fam == "y_inverse_deriv":
        return 3 * S                     # 2D: read 1, write 2
    if fam == "x_advect":
        nin = 3 if "advect3" in name else (2 if "advect2" in name else 1)
        return (nin + 1) * S + nin * R   # read nin spectral rows + nin flows, write 1
    if fam == "x_advect_stage":          # 1D
        return 2 * S + R + (2 * S if stage == 0 else 3 * S)
    if fam == "y_forward":
        return 2 * S
    # RK4 state traffic of the FP64 form: stage 0 reads sol, writes acc; 1 reads sol, acc,
    # writes stg, acc; 2 reads stg, sol, acc, writes stg, sol; 3 reads stg, sol, writes sol.
    # ETDRK4: 0 and 1 read sol, E2, zeta and write the next input and N; 2 reads sol, sol1,
    # N1, N2, E2, zeta, E, alpha, beta and writes sol2 and the partial update; 3 reads the
    # partial update and Gamma and writes sol.
    # The mixed form adds the low parts of its compensated pairs: ETDRK4 reads sol_lo in
    # stages 0-2, writes acc's low part in stage 2 and reads it in stage 3, writes sol_lo
    # in stage 3; RK4 stage 3 reads acc and sol_lo and writes sol_lo.
    mixed = cb == 8
    if stepper.upper().endswith("ETDRK4"):
        state = {0: 3 * S + 2 * C, 1: 3 * S + 2 * C, 2: 6 * S + 5 * C, 3: 2 * S + C}[stage]
        if mixed:
            state += {0: S, 1: S, 2: 2 * S, 3: 2 * S}[stage]
    else:
        state = {0: 2 * S, 1: 4 * S, 2: 5 * S, 3: 3 * S}[stage]
        if mixed and stage == 3:
            state += 3 * S
    if fam in ("z_forward_stage", "y_forward_stage"):
        return S + state
    if fam in ("z_forward_stage_inverse", "y_forward_stage_inverse"):
        # as above, then the next stage's first inverse pass from registers: write 2
        return S + state + 2 * S
    if fam in ("prologue_z_inverse_deriv", "prologue_y_inverse_deriv"):
        return 3 * S
    if fam == "flow_y_synthesis":          # psi0 (FP64 complex) and omega (FP64) in, 2 out
        return nS * 16 + nS * 8 + 2 * S
    if fam.startswith("flow_x_c2r"):
        return S + R
    return None


# kernel-kind name -> (template, MODE, DIR or NIN, epilogue family) of the HIP kernel it
# instantiates; the stage kernels are compiled per epilogue family (1: RK4, 2: ETDRK4 and
# the plain store, 3: all of them, which the FP64 RK4 kinds use), the other column kernels
# carry family 0
_KIND_TEMPLATE = {
    "col_fft_fwd": ("col_kernel", 0, -1, 0), "col_fft_inv": ("col_kernel", 0, 1, 0),
    "col_inv_deriv1": ("col_kernel", 1, 1, 0), "col_inv_deriv2": ("col_kernel", 2, 1, 0),
    "col_fwd_stage": ("col_kernel", 3, -1, 2), "col_flow_synth": ("col_kernel", 4, 1, 0),
    "col_fwd_stage_inv": ("col_kernel", 5, -1, 2),
    "col_fwd_stage_rk4": ("col_kernel", 3, -1, 1), "col_fwd_stage_inv_rk4": ("col_kernel", 5, -1, 1),
    "row_c2r": ("row_kernel", 0, 1, None), "row_r2c": ("row_kernel", 1, 1, None),
    "row_advect1": ("row_kernel", 2, 1, None), "row_advect2": ("row_kernel", 2, 2, None),
    "row_advect3": ("row_kernel", 2, 3, None),
}


def kernel_template(variant: str) -> str:
    """'col_fwd_stage_inv_rk4 N=256 R=8 W=2 occ=3 f32' ->
    'col_kernel<float, 256, 8, 2, 5, -1, 3, 1>', the kernel name the profiler reports for
    that variant."""
    parts = variant.split()
    kv = dict(p.split("=") for p in parts if "=" in p)
    tmpl, mode, last, epf = _KIND_TEMPLATE[parts[0]]
    ty = "double" if parts[-1] == "f64" else "float"
    occ = kv.get("occ", "0")
    if epf == 1 and ty == "double":
        epf = 3
    if "par" in parts:
        if tmpl == "col_kernel":   # pipelined persistent stage kernel
            return f"col_pipe_kernel<{ty}, {kv['N']}, {kv['R']}, {kv['W']}, {occ}, {epf}>"
        # input-parallel advection: row_adv_par_kernel<T, NX, R, NL, NIN, OCC>
        return f"row_adv_par_kernel<{ty}, {kv['N']}, {kv['R']}, {kv['W']}, {last}, {occ}>"
    if tmpl == "col_kernel":
        return f"{tmpl}<{ty}, {kv['N']}, {kv['R']}, {kv['W']}, {mode}, {last}, {occ}, {epf}>"
    return f"{tmpl}<{ty}, {kv['N']}, {kv['R']}, {kv['W']}, {mode}, {last}, {occ}>"


def _counter_rows(d, order=False):
    """(kernel name, counter, value) of every dispatch in a rocprofv3 --pmc csv output folder;
    with order=True (dispatch id, kernel name, counter, value), sorted by dispatch."""
    import csv
    import glob
    out = []
    for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                if order:
                    out.append((int(r["Dispatch_Id"]), r["Kernel_Name"], r["Counter_Name"],
                                float(r["Counter_Value"])))
                else:
                    out.append((r["Kernel_Name"], r["Counter_Name"], float(r["Counter_Value"])))
    return sorted(out) if order else out


def _per_launch_avg(seq, templates):
    """Average of a family's per-dispatch values, launch position by launch position.

    seq: (kernel name, value) of the dispatches of the family's kernels in dispatch order;
    templates: the kernel of each launch of a step, in order.  The dispatches cycle through
    the step's launches; each position is averaged over its dispatches (its own variant),
    then the positions are averaged: the per-launch mean of the step, comparable with the
    family's average algorithmic bytes.  None when the dispatches do not follow that cycle."""
    L = len(templates)
    ds = [(n, x) for (n, x) in seq if any(t in n for t in templates)]
    if not ds or len(ds) % L:
        ds = ds[len(ds) % L:] if ds else ds
    if not ds:
        return None, 0
    pos = [[] for _ in range(L)]
    for i, (n, x) in enumerate(ds):
        if templates[i % L] not in n:
            return None, 0
        pos[i % L].append(x)
    return sum(sum(p) / len(p) for p in pos) / L, len(ds)


def pmc_summary(fetch_dir, write_dir, cal_fetch_dir, cal_write_dir, bench_json):
    """HBM bytes per launch of each kernel family of the headline run, from rocprofv3
    FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each scaled by what the same counter
    reads for a plain 1 GiB device copy (bench.py --copy-calibration)."""
    def avg(rows, counter, pred):
        v = [x for (k, c, x) in rows if c == counter and pred(k)]
        return (sum(v) / len(v), len(v)) if v else (None, 0)
    calib = {}
    for counter, d in (("FETCH_SIZE", cal_fetch_dir), ("WRITE_SIZE", cal_write_dir)):
        kib, n = avg(_counter_rows(d), counter, lambda k: "copyBuffer" in k)
        calib[counter] = dict(kib_per_launch=kib, launches=n,
                              factor=(1 << 30) / (kib * 1024.0) if kib else None)
    bj = json.loads(open(bench_json).read().strip().splitlines()[-1])
    rows = {c: _counter_rows(d, order=True) for c, d in (("FETCH_SIZE", fetch_dir),
                                                          ("WRITE_SIZE", write_dir))}
    fams = {}
    for fam, k in bj["kernels"].items():
        templates = k.get("kernel_per_launch") or [kernel_template(k["kernel"])]
        ent = dict(kernel=k["kernel"], kernel_per_launch=templates,
                   algorithmic_bytes_per_launch=k.get("bytes_per_launch"))
        tot = 0.0
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            seq = [(name, x) for (_, name, c, x) in rows[counter] if c == counter]
            kib, n = _per_launch_avg(seq, templates)
            fac = calib[counter]["factor"]
            ent[counter] = dict(kib_per_launch=kib, launches=n)
            if kib is None or fac is None:
                tot = None
            elif tot is not None:
                tot += kib * 1024.0 * fac
        ent["hbm_bytes_per_launch"] = tot
        fams[fam] = ent
    return dict(calibration=calib, families=fams,
                per_launch_bytes={f: e["hbm_bytes_per_launch"] for f, e in fams.items()})


def trace_summary(trace_dir, bench_out, steps, warmup):
    """Markdown summary of a rocprofv3 --kernel-trace --output-format csv run of bench.py:
    for each kernel family of the FP64 headline, the average duration of the chosen
    variant over the dispatches of the running steps (warmup + timed + profiled steps, the
    last dispatches of that kernel in the FP64 part of the run), beside bench.py's own
    HIP-event figure."""
    import csv
    import glob
    bj = json.loads(open(bench_out).read().strip().splitlines()[-1])
    disp = []
    for path in glob.glob(os.path.join(trace_dir, "**", "*kernel_trace.csv"), recursive=True):
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                disp.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    disp.sort()
    lines = ["| family | kernel | launches/step | rocprof avg us (in-step dispatches) | n | "
             "bench.py HIP-event avg us | ratio |", "|---|---|---|---|---|---|---|"]
    for fam, k in bj["kernels"].items():
        templates = k.get("kernel_per_launch") or [kernel_template(k["kernel"])]
        nstep = k["launches_per_step"] * (warmup + 2 * steps)
        seq = [(n, e - b) for (b, e, n) in disp if any(t in n for t in templates)]
        seq = seq[-nstep:] if len(seq) >= nstep else seq
        avg, n = _per_launch_avg(seq, templates)
        avg = avg / 1e3 if avg is not None else float("nan")
        ev = k["ms_per_launch"] * 1e3
        shown = "`, `".join(dict.fromkeys(templates))
        lines.append(f"| {fam} | `{shown}` | {k['launches_per_step']} | {avg:.1f} | {n} | "
                     f"{ev:.1f} | {avg / ev:.3f} |")
    return "\n".join(lines)


def ensemble_summary(trace_dir, fetch_dir, write_dir, cal_fetch_dir, cal_write_dir, bench_out):
    """Markdown summary of rocprofv3 runs of `bench.py --profile-form ensemble_*`: per kernel
    template, dispatches per ensemble step, average duration (kernel trace) and HBM bytes
    per dispatch (FETCH_SIZE and WRITE_SIZE scaled by the plain-copy calibration), over the
    last steps of the run (the timed ones)."""
    import csv
    import glob
    bj = json.loads(open(bench_out).read().strip().splitlines()[-1])
    steps = bj["steps"]
    disp = []
    for path in glob.glob(os.path.join(trace_dir, "**", "*kernel_trace.csv"), recursive=True):
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                disp.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    disp.sort()
    kern = [d for d in disp if "kernel<" in d[2]]
    # dispatches of one ensemble step: the step kernels of the last `steps` steps
    names = sorted({d[2].split("(")[0] for d in kern})
    per_step = {}
    for n in names:
        ds = [e - b for (b, e, k) in kern if k.split("(")[0] == n]
        per_step[n] = ds
    calib = {}
    for counter, d in (("FETCH_SIZE", cal_fetch_dir), ("WRITE_SIZE", cal_write_dir)):
        v = [x for (k, c, x) in _counter_rows(d) if c == counter and "copyBuffer" in k]
        calib[counter] = (1 << 30) / (sum(v) / len(v) * 1024.0) if v else None
    pmc = {}
    for counter, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        for (k, c, x) in _counter_rows(d):
            if c == counter:
                pmc.setdefault((k.split("(")[0], counter), []).append(x)
    lines = [f"ensemble step: {bj['ms_per_step']:.4f} ms ({bj['value']:.4g} grid-point-steps/s, "
             f"{bj['points_per_step']} points per step)", "",
             "| kernel | dispatches (run) | avg us (last dispatches) | HBM MB / dispatch | GB/s |",
             "|---|---|---|---|---|"]
    tot_us = 0.0
    for n in names:
        ds = per_step[n]
        last = ds[-max(1, len(ds) // 2):]
        avg = sum(last) / len(last) / 1e3
        mb = None
        if calib["FETCH_SIZE"] and calib["WRITE_SIZE"]:
            f = pmc.get((n, "FETCH_SIZE"))
            w = pmc.get((n, "WRITE_SIZE"))
            if f and w:
                mb = (sum(f) / len(f) * 1024 * calib["FETCH_SIZE"] +
                      sum(w) / len(w) * 1024 * calib["WRITE_SIZE"]) / 1e6
        lines.append(f"| `{n}` | {len(ds)} | {avg:.1f} | {mb:.1f} | {mb / avg * 1e3:.0f} |"
                     if mb else f"| `{n}` | {len(ds)} | {avg:.1f} | - | - |")
        tot_us += avg
    return "\n".join(lines)


def TAD_mass(p):
    import passivetracerflows as ptf
    return ptf.TracerAdvectionDiffusion.mass(p)


def TAD_variance(p):
    import passivetracerflows as ptf
    return ptf.TracerAdvectionDiffusion.variance(p)


def gps(points, steps, seconds):
    return points * steps / seconds


def sync():
    """hipDeviceSynchronize through the library, so the timed process binds only the HIP
    runtime the library was built against (importing torch would bind torch's own copy)."""
    from passivetracerflows._lib import call
    call("ptf_device_synchronize")


# ensemble warm-up steps: one block of 16 steps and one single step, so that both captured
# step graphs exist before a timed region
ENS_WARM = 17


def time_steps(stepfn, steps):
    sync()
    t0 = time.perf_counter()
    stepfn(steps)
    sync()
    return time.perf_counter() - t0


# tuning keys already retuned by this process: a later problem of the same workload (the
# full-length error runs, the repeats) takes the fresh stored choice instead of tuning
# again, so that the stored variants are the ones the headline was timed with
_RETUNED = set()


def _tune(obj, autotune):
    """autotune: False (default variants), "cache" (stored choice, else tune and store) or
    "retune" (tune and overwrite the stored choice, once per key and process)."""
    if not autotune:
        return "default"
    if autotune == "retune":
        from passivetracerflows import tuning
        key = obj.tuning_key() if hasattr(obj, "tuning_key") else _problem_key(obj)
        if key not in _RETUNED:
            _RETUNED.add(key)
            data = tuning.load(TUNING)
            data.pop(key, None)
            tuning.store(TUNING, data)
    return obj.autotune(cache=TUNING)


def _problem_key(p):
    from passivetracerflows.fourierflows import tuning_key
    return tuning_key(p)


def headline(dev, steps, warmup, T=np.float64, autotune="cache"):
    import passivetracerflows as ptf
    from passivetracerflows import workloads as W
    TAD = ptf.TracerAdvectionDiffusion
    cfg = W.abc_3d(256)
    u, v, w = cfg["flow"]
    prob = TAD.Problem(dev, ptf.ThreeDAdvectingFlow(u=lambda x, y, z: u, v=lambda x, y, z: v,
                                                     w=lambda x, y, z: w),
                       nx=256, κ=cfg["kappa"], dt=cfg["dt"], stepper="RK4", T=T)
    TAD.set_c_(prob, cfg["c0"])
    prob.tuned = _tune(prob, autotune)
    m0 = TAD.mass(prob)
    v0 = TAD.variance(prob)
    ptf.stepforward_(prob, warmup)
    secs = time_steps(lambda k: ptf.stepforward_(prob, k), steps)
    m1 = TAD.mass(prob)
    v1 = TAD.variance(prob)
    return cfg, prob, secs, (m0, m1, v0, v1)


def repeat_and_pcie(prob, cfg, steps):
    """Two more timed segments of `steps` steps (run-to-run spread), and one segment
    that includes the host boundary: upload of u, v, w and c0 (ptf_set_flow, set_c!),
    the steps, and the download of c (updatevars!)."""
    import passivetracerflows as ptf
    from passivetracerflows import traceradvectiondiffusion as tad
    reps = [time_steps(lambda k: ptf.stepforward_(prob, k), steps) for _ in range(2)]
    sync()
    t0 = time.perf_counter()
    tad._upload_fields(prob._h, list(cfg["flow"]))
    tad.set_c_(prob, cfg["c0"])
    ptf.stepforward_(prob, steps)
    tad.updatevars_(prob)
    sync()
    pcie = time.perf_counter() - t0
    return reps, pcie


def slot_table(prob, nS, nR, cb, rb, steps, stepper="RK4"):
    """Per-launch durations measured inside `steps` running steps (a HIP event around
    every launch on the problem's stream), with each launch's algorithmic bytes."""
    rows = []
    for i, (name, ms) in enumerate(zip(prob.slots(), prob.step_profiled(steps))):
        stage = int(name[5]) - 1 if name.startswith("stage") else 0
        b = kernel_bytes(name, nS, nR, cb, rb, stage, stepper)
        rows.append(dict(slot=i, name=name, ms=ms, bytes=b,
                         gbs=(b / (ms * 1e-3) / 1e9) if b else None))
    return rows


def dominant(rows):
    """Kernel families of a step (a slot name without its stage prefix): summed in-step time,
    launches, algorithmic bytes, and the variant of every launch in step order (the stage
    kernels are tuned per stage, so a family can mix variants); `kernel` names them."""
    fams = {}
    for r in rows:
        key = r["name"].split(" [")[0]
        key = key.split("_", 1)[1] if key.startswith("stage") else key
        f = fams.setdefault(key, dict(total_ms=0.0, n=0, bytes=0, per_launch=[]))
        f["total_ms"] += r["ms"]
        f["n"] += 1
        f["bytes"] += r["bytes"] or 0
        f["per_launch"].append(r["name"].split(" [")[1][:-1])
    for f in fams.values():
        v = f["per_launch"]
        if len(set(v)) == 1:
            f["kernel"] = v[0]
        else:
            f["kernel"] = "; ".join(
                f"{u} (launch {', '.join(str(i + 1) for i, w in enumerate(v) if w == u)})"
                for u in dict.fromkeys(v))
        f["templates"] = [kernel_template(u) for u in v]
    name = max(fams, key=lambda k: fams[k]["total_ms"])
    f = fams[name]
    avg_ms = f["total_ms"] / f["n"]
    avg_bytes = f["bytes"] / f["n"]
    return name, f, avg_ms, avg_bytes, fams


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cfg, steps):
    from oracle import tad as O
    cores = min(16, os.cpu_count() or 1)
    O.set_workers(cores)
    g = O.Grid(3, 256)
    p = O.Problem(g, cfg["flow"], kappa=cfg["kappa"], dt=cfg["dt"], stepper="RK4")
    p.set_c(cfg["c0"])
    p.stepforward(1)  # warm-up
    t0 = time.perf_counter()
    p.stepforward(steps)
    dt = time.perf_counter() - t0
    return dict(value=gps(g.npoints, steps, dt), unit="grid-point-steps/s", cores=cores,
                kind="port", cpu=cpu_model(), nproc=os.cpu_count(),
                sample=f"FP64 oracle restatement (scipy.fft pocketfft, workers={cores}; numpy "
                       f"elementwise single-threaded), 3D ABC 256^3 RK4, {steps} steps after 1 "
                       f"warm-up step, {dt:.1f} s")


def ensemble_parity(members, e32, e64, e64r, dt, steps):
    """Every member's FP32 state against its own FP64 run after `steps` steps, with the
    FP64 problem's own sensitivity as the control: e64r is the FP64 ensemble run from the
    inputs (c0, flow) rounded to FP32, and spread = relL2(e64r, e64) per member.  Members
    are classed by workloads.member_class (RK4 diffusion limit, spectral resolution);
    `diverged` are members whose FP64 run leaves the FP32 range."""
    from passivetracerflows import workloads as W
    classes = {"resolved": 0, "under_resolved": 0, "diffusion_unstable": 0}
    worst_resolved, within, within_ctrl, diverged = 0.0, 0, 0, 0
    ratio_resolved = 0.0
    for i, m in enumerate(members):
        a, b = e32.get_sol(i), e64.get_sol(i)
        cls = W.member_class(m, dt, b)
        classes[cls] += 1
        with np.errstate(all="ignore"):
            if not (np.all(np.isfinite(b)) and np.abs(b).max() < 1e30):
                diverged += 1
                continue
            r = float(np.linalg.norm(a - b) / np.linalg.norm(b))
            sp = float(np.linalg.norm(e64r.get_sol(i) - b) / np.linalg.norm(b))
        if cls == "resolved":
            worst_resolved = max(worst_resolved, r)
            ratio_resolved = max(ratio_resolved, r / sp if sp > 0 else 0.0)
        within += int(r <= 1e-5)
        within_ctrl += int(r <= max(1e-5, 300 * sp))
    return dict(steps=steps, members=len(members), classes=classes, diverged=diverged,
                max_relL2_resolved=worst_resolved, members_within_1e5=within,
                max_error_over_spread_resolved=ratio_resolved,
                members_within_max_1e5_300_spread=within_ctrl)


def mixed_errors(TAD, p32, p64, c0):
    """Mixed form against the FP64 run of the same config after the same steps: relative
    L2 of sol, mass error relative to sum|c0|, variance error relative to the variance."""
    s32, s64 = p32.sol, p64.sol
    v64 = TAD.variance(p64)
    return dict(relL2_sol=float(np.linalg.norm(s32 - s64) / np.linalg.norm(s64)),
                mass_err_rel_l1=abs(TAD.mass(p32) - TAD.mass(p64)) / float(np.abs(c0).sum()),
                variance_err_rel=abs(TAD.variance(p32) - v64) / v64)


def full_length_mixed_errors(dev, autotune="cache"):
    """The mixed form against the FP64 form after each workload's configured step count
    (3D 256^3 x 500, 2D 256^2 x 2000, 2D 2048^2 ETDRK4 x 1000): relative L2 of sol, mass
    error relative to sum|c0|, variance error relative to the variance."""
    import passivetracerflows as ptf
    from passivetracerflows import workloads as W
    TAD = ptf.TracerAdvectionDiffusion
    out = {}
    cases = (
        ("abc3d_256_rk4", W.abc_3d(256),
         lambda cfg, T: TAD.Problem(dev, ptf.ThreeDAdvectingFlow(*[(lambda *a, f=f: f) for f in cfg["flow"]]),
                                    nx=256, κ=cfg["kappa"], dt=cfg["dt"], T=T)),
        ("cellular2d_256_rk4", W.cellular_2d(256),
         lambda cfg, T: TAD.Problem(dev, ptf.TwoDAdvectingFlow(*[(lambda *a, f=f: f) for f in cfg["flow"]]),
                                    nx=256, κ=cfg["kappa"], dt=cfg["dt"], T=T)),
        ("turbulent2d_2048_etdrk4", W.turbulent_2d(2048),
         lambda cfg, T: TAD.Problem(dev, ptf.SpectralStreamfunctionFlow(cfg["psih0"], cfg["omega"]),
                                    nx=2048, κ=0.0, η=0.0, κh=cfg["kappah"], nκh=cfg["nkappah"],
                                    dt=cfg["dt"], stepper="ETDRK4", T=T)),
    )
    for name, cfg, make in cases:
        probs = {}
        for T in (np.float64, np.float32):
            p = make(cfg, T)
            TAD.set_c_(p, cfg["c0"])
            _tune(p, autotune)
            ptf.stepforward_(p, cfg["nsteps"])
            probs[T] = p
        out[name] = dict(steps=cfg["nsteps"], **mixed_errors(TAD, probs[np.float32], probs[np.float64],
                                                            cfg["c0"]))
        del probs
    return out


def extra_forms(dev, steps, ens_steps, autotune="cache"):
    import passivetracerflows as ptf
    from passivetracerflows import workloads as W
    TAD = ptf.TracerAdvectionDiffusion
    out = {}
    # 2D cellular 256^2, FP64 and mixed
    cfg = W.cellular_2d(256)
    u, v = cfg["flow"]
    probs = {}
    for T, tag in ((np.float64, "fp64"), (np.float32, "mixed")):
        p = TAD.Problem(dev, ptf.TwoDAdvectingFlow(u=lambda x, y: u, v=lambda x, y: v), nx=256,
                        κ=cfg["kappa"], dt=cfg["dt"], T=T)
        TAD.set_c_(p, cfg["c0"])
        _tune(p, autotune)
        ptf.stepforward_(p, 10)
        k = max(steps * 4, 200)
        s = time_steps(lambda n: ptf.stepforward_(p, n), k)
        out[f"cellular2d_256_{tag}"] = dict(value=gps(256 * 256, k, s), ms_per_step=1e3 * s / k,
                                            steps=k)
        probs[tag] = p
    out["cellular2d_256_mixed"]["vs_fp64_after_steps"] = dict(
        steps=10 + max(steps * 4, 200), **mixed_errors(TAD, probs["mixed"], probs["fp64"], cfg["c0"]))
    # 2D 2048^2 random-phase flow, ETDRK4
    cfg = W.turbulent_2d(2048)
    for T, tag in ((np.float64, "fp64"), (np.float32, "mixed")):
        p = TAD.Problem(dev, ptf.SpectralStreamfunctionFlow(cfg["psih0"], cfg["omega"]), nx=2048,
                        κ=0.0, η=0.0, κh=cfg["kappah"], nκh=cfg["nkappah"], dt=cfg["dt"],
                        stepper="ETDRK4", T=T)
        TAD.set_c_(p, cfg["c0"])
        _tune(p, autotune)
        ptf.stepforward_(p, 3)
        k = max(steps // 2, 20)
        s = time_steps(lambda n: ptf.stepforward_(p, n), k)
        out[f"turbulent2d_2048_etdrk4_{tag}"] = dict(value=gps(2048 * 2048, k, s),
                                                     ms_per_step=1e3 * s / k, steps=k)
        probs[tag] = p
    out["turbulent2d_2048_etdrk4_mixed"]["vs_fp64_after_steps"] = dict(
        steps=3 + max(steps // 2, 20), **mixed_errors(TAD, probs["mixed"], probs["fp64"], cfg["c0"]))
    # 512-member ensemble, FP32 bulk, hipGraph; parity vs the FP64 ensemble on the GPU
    ens_cfg = W.ensemble_2d()
    members = [dict(m, **W.ensemble_member(m["n"], m["seed"])) for m in ens_cfg["members"]]
    npts = sum(m["n"] ** 2 for m in members)
    res = {}
    for prec in ("mixed", "fp64"):
        e = ptf.Ensemble([dict(n=m["n"], kappa=m["kappa"]) for m in members], dt=ens_cfg["dt"],
                         precision=prec)
        for i, m in enumerate(members):
            e.set_flow(i, *m["flow"])
            e.set_c(i, m["c0"])
        _tune(e, autotune)
        # both step graphs (a block of 16 steps, a single step) captured and replayed once
        # outside the timed region
        e.step(ENS_WARM, graph=True)
        k = min(ens_steps, ens_cfg["nsteps"]) - ENS_WARM
        s = time_steps(lambda n: e.step(n, graph=True), k)
        res[prec] = (e, s, k)
    e32, s32, k = res["mixed"]
    e64, s64, _ = res["fp64"]
    # the control: FP64 from inputs rounded to FP32 (not timed)
    r32 = lambda a: np.asarray(a, np.float32).astype(np.float64)
    e64r = ptf.Ensemble([dict(n=m["n"], kappa=m["kappa"]) for m in members], dt=ens_cfg["dt"],
                        precision="fp64")
    for i, m in enumerate(members):
        e64r.set_flow(i, *[r32(f) for f in m["flow"]])
        e64r.set_c(i, r32(m["c0"]))
    _tune(e64r, autotune)   # the same stored variants as the FP64 run it is compared with
    e64r.step(k + ENS_WARM, graph=True)
    out["ensemble512_fp32_graph"] = dict(value=gps(npts, k, s32), ms_per_step=1e3 * s32 / k, steps=k,
                                         points_per_step=npts)
    out["ensemble512_fp64_graph"] = dict(value=gps(npts, k, s64), ms_per_step=1e3 * s64 / k, steps=k)
    out["ensemble512_parity"] = ensemble_parity(members, e32, e64, e64r, ens_cfg["dt"], k + ENS_WARM)
    return out


PROFILE_FORMS = ("abc3d_fp64", "abc3d_mixed", "turb2d_fp64", "turb2d_mixed", "ensemble_fp32",
                 "ensemble_fp64")


def profile_form(dev, form, steps, warmup, autotune):
    """One workload alone (for a rocprofv3 run): stored variants, `warmup` steps, `steps`
    timed steps, then (single problems) `steps` steps with an event around every launch.
    Prints one JSON line: ms/step, GPS/s and, per kernel family, the variant, its
    average in-step duration and its algorithmic bytes."""
    import passivetracerflows as ptf
    from passivetracerflows import workloads as W
    TAD = ptf.TracerAdvectionDiffusion
    out = dict(form=form, steps=steps, warmup=warmup)
    if form.startswith("ensemble"):
        cfg = W.ensemble_2d()
        members = [dict(m, **W.ensemble_member(m["n"], m["seed"])) for m in cfg["members"]]
        prec = "mixed" if form.endswith("fp32") else "fp64"
        e = ptf.Ensemble([dict(n=m["n"], kappa=m["kappa"]) for m in members], dt=cfg["dt"],
                         precision=prec)
        for i, m in enumerate(members):
            e.set_flow(i, *m["flow"])
            e.set_c(i, m["c0"])
        out["variants"] = _tune(e, autotune)
        e.step(max(warmup, ENS_WARM), graph=True)
        secs = time_steps(lambda n: e.step(n, graph=True), steps)
        npts = sum(m["n"] ** 2 for m in members)
        out.update(ms_per_step=1e3 * secs / steps, value=gps(npts, steps, secs),
                   points_per_step=npts, slots=e.slots())
        print(json.dumps(out), flush=True)
        return
    T = np.float64 if form.endswith("fp64") else np.float32
    if form.startswith("abc3d"):
        cfg, prob, secs, _ = headline(dev, steps, warmup, T=T, autotune=autotune)
        n, stepper = (256, 256, 256), "RK4"
    else:
        cfg = W.turbulent_2d(2048)
        prob = TAD.Problem(dev, ptf.SpectralStreamfunctionFlow(cfg["psih0"], cfg["omega"]), nx=2048,
                           κ=0.0, η=0.0, κh=cfg["kappah"], nκh=cfg["nkappah"], dt=cfg["dt"],
                           stepper="ETDRK4", T=T)
        TAD.set_c_(prob, cfg["c0"])
        prob.tuned = _tune(prob, autotune)
        ptf.stepforward_(prob, warmup)
        secs = time_steps(lambda k: ptf.stepforward_(prob, k), steps)
        n, stepper = (2048, 2048, 1), "ETDRK4"
    npts = n[0] * n[1] * n[2]
    nS = (n[0] // 2 + 1) * n[1] * n[2]
    cb, rb = (16, 8) if T is np.float64 else (8, 4)
    rows = slot_table(prob, nS, npts, cb, rb, steps, stepper)
    _, _, _, _, fams = dominant(rows)
    out.update(ms_per_step=1e3 * secs / steps, value=gps(npts, steps, secs), variants=prob.tuned,
               kernels={k: dict(kernel=f["kernel"], kernel_per_launch=f["templates"],
                                ms_per_launch=f["total_ms"] / f["n"],
                                launches_per_step=f["n"], bytes_per_launch=f["bytes"] / f["n"],
                                GBs=(f["bytes"] / f["total_ms"] * 1e3 / 1e9) if f["bytes"] else None)
                        for k, f in fams.items()})
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--retune", action="store_true",
                    help="run the autotuner and overwrite its stored choices "
                         "(default: apply the stored choices, tuning only what is missing)")
    ap.add_argument("--ensemble-steps", type=int, default=1000)
    ap.add_argument("--pmc-summary", nargs=5,
                    metavar=("FETCH_DIR", "WRITE_DIR", "CAL_FETCH_DIR", "CAL_WRITE_DIR", "BENCH_OUT"),
                    help="summarise rocprofv3 --pmc csv outputs into per-launch HBM bytes (JSON) "
                         "and exit; runs on the CPU")
    ap.add_argument("--ensemble-summary", nargs=6,
                    metavar=("TRACE_DIR", "FETCH_DIR", "WRITE_DIR", "CAL_FETCH_DIR", "CAL_WRITE_DIR",
                             "BENCH_OUT"),
                    help="summarise rocprofv3 kernel-trace and --pmc outputs of a --profile-form "
                         "ensemble run, and exit")
    ap.add_argument("--trace-summary", nargs=2, metavar=("TRACE_DIR", "BENCH_OUT"),
                    help="summarise a rocprofv3 --kernel-trace csv output of this script (run "
                         "with the same --steps/--warmup) against its JSON line, and exit")
    ap.add_argument("--profile-form", choices=PROFILE_FORMS,
                    help="run one workload alone with the stored variants (for a profiler run) "
                         "and print its per-kernel timings; no headline line")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of 1 GiB five times (the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE) and exit")
    args = ap.parse_args()
    faulthandler.enable()
    if args.pmc_summary:
        print(json.dumps(pmc_summary(*args.pmc_summary), indent=1))
        return
    if args.trace_summary:
        print(trace_summary(*args.trace_summary, args.steps, args.warmup))
        return
    if args.ensemble_summary:
        print(ensemble_summary(*args.ensemble_summary))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs a single problem (or ensemble) per GPU "
              "and has no multi-GPU path; use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        import torch          # the runtime's plain device copy, as the HBM-counter yardstick
        n = 1 << 27                              # 2^27 doubles = 1 GiB
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": 8 * n, "copy_calibration_bytes_written": 8 * n,
                          "launches": 5}), flush=True)
        return
    import passivetracerflows as ptf
    dev = ptf.ROCm(0)
    autotune = False if args.no_autotune else ("retune" if args.retune else "cache")
    if args.profile_form:
        profile_form(dev, args.profile_form, args.steps, args.warmup, autotune)
        return

    cfg, prob, secs, (m0, m1, v0, v1) = headline(dev, args.steps, args.warmup, autotune=autotune)
    s64 = prob.sol            # state after warmup + steps, for the mixed-form comparison
    reps, pcie_secs = repeat_and_pcie(prob, cfg, args.steps)
    l1 = float(np.abs(cfg["c0"]).sum())
    npts = 256 ** 3
    value = gps(npts, args.steps, secs)
    nS, nR = 129 * 256 * 256, npts
    rows = slot_table(prob, nS, nR, 16, 8, args.steps)
    dname, fam, avg_ms, avg_bytes, fams = dominant(rows)
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
    step_bytes = sum(r["bytes"] for r in rows if r["bytes"])
    step_ms_kernels = sum(r["ms"] for r in rows)

    # measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a plain
    # copy), used only when the measured kernel is the variant timed here
    traffic, traffic_variant = None, None
    tpath = os.path.join(ROOT, "profiles", "traffic_abc3d_256_fp64.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        ent = tj.get("families", {}).get(dname, {})
        traffic_variant = ent.get("kernel_per_launch") or ent.get("kernel")
        if ent.get("kernel_per_launch") == fam["templates"]:
            traffic = ent.get("hbm_bytes_per_launch")

    # mixed form of the same workload + its error against the FP64 run
    _, pm, secs32, _ = headline(dev, args.steps, args.warmup, T=np.float32, autotune=autotune)
    s32 = pm.sol
    mixed_err = float(np.linalg.norm(s32 - s64) / np.linalg.norm(s64))
    mixed_diag = dict(mass_err_rel_l1=abs(TAD_mass(pm) - m1) / float(np.abs(cfg["c0"]).sum()),
                      variance_err_rel=abs(TAD_variance(pm) - v1) / v1)

    forms = {
        "abc3d_256_fp64": dict(value=value, ms_per_step=1e3 * secs / args.steps, steps=args.steps,
                               kernel_ms_per_step=step_ms_kernels,
                               schedule_bytes_per_step=step_bytes,
                               schedule_GBs=step_bytes / (secs / args.steps) / 1e9,
                               pass_model_GBs=PASS_MODEL_B_PER_GPS * value / 1e9,
                               pass_model_frac=PASS_MODEL_B_PER_GPS * value / 1e9 / HBM_PEAK_GBS,
                               mass_drift_rel_to_l1=abs(m1 - m0) / l1,
                               variance_ratio=v1 / v0,
                               repeat_ms_per_step=[1e3 * r / args.steps for r in reps],
                               median_value=float(np.median([gps(npts, args.steps, r)
                                                             for r in [secs] + reps]))),
        "abc3d_256_fp64_pcie_inclusive": dict(
            value=gps(npts, args.steps, pcie_secs), ms_per_step=1e3 * pcie_secs / args.steps,
            steps=args.steps, includes="ptf_set_flow(u,v,w) + set_c! + steps + updatevars!"),
        "abc3d_256_mixed": dict(value=gps(npts, args.steps, secs32),
                                ms_per_step=1e3 * secs32 / args.steps,
                                relL2_vs_fp64_after_steps=mixed_err, **mixed_diag,
                                steps_compared=args.steps + args.warmup),
    }
    if not args.no_extra:
        forms.update(extra_forms(dev, args.steps, args.ensemble_steps, autotune))
        forms["mixed_vs_fp64_full_length"] = full_length_mixed_errors(dev, autotune)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cfg, args.cpu_steps)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "grid-point-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * secs / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded ABC flow and random-phase k^-2 initial field)",
        "config": {"workload": "3D steady ABC flow, nx=ny=nz=256, L=2pi, kappa=1e-3, RK4, "
                               "dt=5e-3, FP64 plain form", "grid": [256, 256, 256],
                   "stepper": "RK4", "points": npts},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_variant": traffic_variant,
                     "kernel": f"{dname} [{fam['kernel']}]",
                     "kernel_template": fam["templates"], "variants": prob.tuned,
                     "avg_launch_ms": avg_ms,
                     "algorithmic_bytes_per_launch": avg_bytes},
        "cpu_baseline": cpu,
        "kernels": {k: dict(kernel=f["kernel"], kernel_per_launch=f["templates"],
                            ms_per_launch=f["total_ms"] / f["n"],
                            launches_per_step=f["n"], bytes_per_launch=f["bytes"] / f["n"],
                            GBs=(f["bytes"] / f["total_ms"] * 1e3 / 1e9) if f["bytes"] else None)
                    for k, f in fams.items()},
        "forms": forms,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
