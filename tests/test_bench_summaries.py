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
        v = "col_fwd_stage_inv N=256 R=8 W=8 occ=3 wg f64" if s != 2 else \
            "col_fwd_stage_inv N=256 R=8 W=8 wg f64"
        rows.append(dict(name=f"stage{s + 1}_z_forward_stage_inverse [{v}]", ms=0.2, bytes=8e8))
        rows.append(dict(name=f"stage{s + 1}_x_advect [row_advect3 N=256 R=4 W=8 occ=3 f64]",
                         ms=0.1, bytes=9e8))
    name, fam, avg_ms, avg_bytes, fams = bench.dominant(rows)
    assert name == "z_forward_stage_inverse"
    assert fam["templates"] == ["col_kernel<double, 256, 8, 8, 5, -1, 3>"] * 2 + \
        ["col_kernel<double, 256, 8, 8, 5, -1, 0>", "col_kernel<double, 256, 8, 8, 5, -1, 3>"]
    assert "(launch 1, 2, 4)" in fam["kernel"] and "(launch 3)" in fam["kernel"]
    assert fams["x_advect"]["kernel"] == "row_advect3 N=256 R=4 W=8 occ=3 f64"
    assert abs(avg_ms - 0.2) < 1e-12 and abs(avg_bytes - 8e8) < 1e-3
