"""bench.py's per-family summaries when a family's launches use different kernel variants
(the stage kernels are tuned per stage), on synthetic dispatch lists (host only)."""
import bench


def test_per_launch_average_follows_the_step_cycle():
    templates = ["kA<1>", "kA<1>", "kB<2>", "kA<1>"]       # stage 3 uses another variant
    seq = []
    for step in range(3):
        for pos, t in enumerate(templates):
            seq.append((f"void {t}(args)", 10.0 * (pos + 1) + step))
    avg, n = bench._per_launch_avg(seq, templates)
    assert n == 12
    # positions average 11, 21, 31, 41 -> 26
    assert abs(avg - 26.0) < 1e-12
    # a leading partial cycle (dispatches before the timed steps) is dropped
    avg2, n2 = bench._per_launch_avg([("void kA<1>(x)", 99.0)] + seq, templates)
    assert n2 == 12 and abs(avg2 - 26.0) < 1e-12
    # dispatches that do not follow the cycle are not attributed
    bad = [(n.replace("kB<2>", "kA<1>"), x) for (n, x) in seq]
    assert bench._per_launch_avg(bad, templates) == (None, 0)


def test_dominant_names_each_variant_with_its_launches():
    rows = []
    for s in range(4):
        v = "col_fwd_stage_inv_rk4 N=256 R=8 W=8 occ=3 wg f64" if s != 2 else \
            "col_fwd_stage_inv_rk4 N=256 R=8 W=8 wg f64"
        rows.append(dict(name=f"stage{s + 1}_z_forward_stage_inverse [{v}]", ms=0.2, bytes=8e8))
        rows.append(dict(name=f"stage{s + 1}_x_advect [row_advect3 N=256 R=4 W=8 occ=3 f64]",
                         ms=0.1, bytes=9e8))
    name, fam, avg_ms, avg_bytes, fams = bench.dominant(rows)
    assert name == "z_forward_stage_inverse"
    assert fam["templates"] == ["col_kernel<double, 256, 8, 8, 5, -1, 3, 3>"] * 2 + \
        ["col_kernel<double, 256, 8, 8, 5, -1, 0, 3>", "col_kernel<double, 256, 8, 8, 5, -1, 3, 3>"]
    assert "(launch 1, 2, 4)" in fam["kernel"] and "(launch 3)" in fam["kernel"]
    assert fams["x_advect"]["kernel"] == "row_advect3 N=256 R=4 W=8 occ=3 f64"
    assert abs(avg_ms - 0.2) < 1e-12 and abs(avg_bytes - 8e8) < 1e-3


def test_kernel_template_names_every_form():
    kt = bench.kernel_template
    assert kt("col_fwd_stage_inv N=2048 R=8 W=1 f64") == "col_kernel<double, 2048, 8, 1, 5, -1, 0, 2>"
    assert kt("col_fwd_stage_rk4 N=64 R=4 W=8 f32") == "col_kernel<float, 64, 4, 8, 3, -1, 0, 1>"
    assert kt("col_fft_fwd N=256 R=4 W=8 f64") == "col_kernel<double, 256, 4, 8, 0, -1, 0, 0>"
    assert kt("col_fwd_stage_inv_rk4 N=256 R=8 W=2 par f32") == \
        "col_pipe_kernel<float, 256, 8, 2, 0, 1>"
    # FP64 RK4 stage kernels are compiled with every epilogue (family 3)
    assert kt("col_fwd_stage_inv_rk4 N=256 R=8 W=8 occ=3 wg f64") == \
        "col_kernel<double, 256, 8, 8, 5, -1, 3, 3>"
    assert kt("row_advect2 N=512 R=4 W=8 occ=3 f32") == "row_kernel<float, 512, 4, 8, 2, 2, 3>"
    assert kt("row_advect3 N=256 R=4 W=8 par occ=3 f32") == \
        "row_adv_par_kernel<float, 256, 4, 8, 3, 3>"


def test_retune_tunes_each_key_once(tmp_path, monkeypatch):
    """bench.py --retune: the first problem of a workload drops the stored choice and tunes;
    later problems of the same workload reuse what it stored."""
    from passivetracerflows import tuning
    path = str(tmp_path / "tuning.json")
    tuning.store(path, {"k": [["slot", 4, 4, False, False, False]], "other": []})
    monkeypatch.setattr(bench, "TUNING", path)
    monkeypatch.setattr(bench, "_RETUNED", set())
    seen = []

    class Obj:
        def tuning_key(self):
            return "k"

        def autotune(self, cache):
            seen.append("k" in tuning.load(cache))
            data = tuning.load(cache)
            data.setdefault("k", [["slot", 8, 8, True, False, True]])
            tuning.store(cache, data)
            return "tuned" if not seen[-1] else "cache"

    assert bench._tune(Obj(), "retune") == "tuned"     # stored choice dropped, tuned
    assert bench._tune(Obj(), "retune") == "cache"     # the fresh choice reused
    assert seen == [False, True]
    assert tuning.load(path)["k"] == [["slot", 8, 8, True, False, True]]
    assert "other" in tuning.load(path)


def test_stage_bytes_count_the_mixed_low_parts():
    """Stage-kernel algorithmic bytes: FP64 RK4 14 S-fields of state per step, the mixed
    form 17 (stage 4 reads acc and sol_lo and writes sol_lo); ETDRK4 mixed adds 6 low-part
    fields per step.  Each fused stage kernel also moves its input and two outputs."""
    nS, nR = 1000, 2000
    name = "stage{}_y_forward_stage_inverse [x]"
    def total(cb, stepper):
        return sum(bench.kernel_bytes(name.format(s + 1), nS, nR, cb, cb // 2, s, stepper)
                   for s in range(4))
    assert total(16, "RK4") == (14 + 4 * 3) * nS * 16
    assert total(8, "RK4") == (17 + 4 * 3) * nS * 8
    C = nS * 8
    assert total(16, "ETDRK4") == (14 + 4 * 3) * nS * 16 + 10 * C
    assert total(8, "ETDRK4") == (20 + 4 * 3) * nS * 8 + 10 * C
