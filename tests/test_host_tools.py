"""Host-side tools that need no GPU: the frozen work model, the synthetic
workload generator, the serving workload helpers, the autotuner's cache and
the benchmark's profile reader."""
import json

import numpy as np
import pytest

from milkyway_client_amd import autotune, serving, synth, workmodel


def test_work_model_is_frozen():
    assert workmodel.flops_per_point(1) == 77
    assert workmodel.flops_per_point(3) == 161
    assert workmodel.flops_per_point(4) == 203
    full = synth.FULL_SIZE.regions()
    assert workmodel.integral_points(full, 120) == 91_392_000_000
    assert workmodel.integral_flops(full, 120, 3) == pytest.approx(14.714112e12)
    ref = synth.REFERENCE_SIZE.regions()
    assert workmodel.integral_flops(ref, 30, 1) == pytest.approx(1.655808e9)
    # 1.088 M columns x 32 B + two regions x (700 x 120 x 16 B + 700 x 64 B)
    assert workmodel.integral_bytes(full, 120) == 34_816_000 + 2 * (1_344_000 + 44_800)


def test_synthetic_workloads_are_seeded_and_in_range():
    p1, s1 = synth.workload(0, synth.REFERENCE_SIZE, 2)
    p2, s2 = synth.workload(0, synth.REFERENCE_SIZE, 2)
    assert np.array_equal(s1, s2) and p1.background_parameters == p2.background_parameters
    assert s1.shape == (10_000, 3)
    r = s1[:, 2]
    assert r.min() >= synth.r_kpc_from_g(16.0) and r.max() <= synth.r_kpc_from_g(22.5)
    q, r0 = p1.background_parameters[1:3]
    assert 0.4 <= q <= 0.9 and 4 <= r0 <= 8
    for s in p1.streams:
        mu, rs, th, ph, sig = s.parameters
        assert 310 <= mu <= 410 and 5 <= rs <= 40 and 0 <= th <= np.pi and 1 <= sig <= 5
        assert -4 <= s.weight <= 0
    v = synth.parameter_vectors(3, p1, 5)
    assert v.shape == (5, p1.optimized_parameter_count())
    assert np.array_equal(v, synth.parameter_vectors(3, p1, 5))


def test_full_size_shape():
    g = synth.FULL_SIZE.regions()
    assert [(x.r_steps, x.mu_steps, x.nu_steps) for x in g] == [(700, 1600, 640), (700, 400, 160)]
    assert g[1].mu_min >= g[0].mu_min and g[1].mu_max <= g[0].mu_max
    assert g[1].mu_max - g[1].mu_min == pytest.approx(25.0)


def test_serving_workload_helpers():
    arr = serving.poisson_arrivals(2000, rate=10.0, seed=1)
    assert np.all(np.diff(arr) > 0)
    assert np.mean(np.diff(arr)) == pytest.approx(0.1, rel=0.1)
    shapes = [synth.WedgeShape(5, 6, 4, 10, 20)]
    wedges = serving.make_wedges(3, seed=2, shapes=shapes)
    assert all(1 <= p.number_streams <= 4 for p, _, _ in wedges)
    reqs = serving.make_requests(wedges, 50, seed=3, fp64_fraction=0.5)
    assert {r.precision for r in reqs} == {0, 1}
    assert all(len(r.vector) == wedges[r.wedge][0].optimized_parameter_count() for r in reqs)


def test_autotune_cache_round_trip(tmp_path):
    path = tmp_path / "t.json"
    assert autotune.load_cache(path) == {}
    autotune.save_cache({"k": {"variant": 2, "blocks_per_cu": 4}}, path)
    assert autotune.load_cache(path)["k"]["variant"] == 2
    sig = autotune.signature([(None, 16, "a"), (None, 2, "b")], 1)
    assert sig == autotune.signature([(None, 2, "b"), (None, 16, "a")], 1)
    assert sig != autotune.signature([(None, 2, "b"), (None, 16, "a")], 0)


def test_bench_reads_measured_traffic():
    import bench
    assert bench.measured_traffic("integral_kernel_fp64_full") > 36_000_000
    assert bench.measured_traffic("no_such_kernel") is None
    assert bench.measured_counters("no_such_kernel") is None


def test_bench_cpu_sample_keeps_the_cells():
    """The CPU baseline's sample keeps each region's first mu rows with the
    same cell size, so that its cost per row is the full wedge's."""
    import bench
    p, _ = synth.workload(1, synth.FULL_SIZE, 3)
    q = bench.first_rows(p, 16)
    for a, b in zip(p.regions, q.regions):
        assert b.mu_steps == 16 and b.mu_min == a.mu_min
        assert (b.mu_max - b.mu_min) / b.mu_steps == pytest.approx((a.mu_max - a.mu_min) / a.mu_steps)
        assert (b.r_steps, b.nu_steps, b.r_min, b.r_max) == (a.r_steps, a.nu_steps, a.r_min, a.r_max)
    assert p.regions[0].mu_steps == 1600          # the original is untouched


def test_bench_vector_sets_are_fresh():
    import bench
    plist = [synth.workload(100 + k, s, 4) for k, s in enumerate(synth.BATCH_SHAPES[:2])]
    a, b = bench.vector_set(0, plist), bench.vector_set(1, plist)
    assert all(x.shape == (16, plist[0][0].optimized_parameter_count()) for x in a)
    assert not any(np.array_equal(x, y) for x, y in zip(a, b))
    assert all(np.array_equal(x, y) for x, y in zip(a, bench.vector_set(0, plist)))


def test_bench_reads_the_headline_counters():
    import bench
    c = bench.measured_counters("integral_persistent_mixed_batch64")
    assert c is not None and c["source"].startswith("profiles/")
    for k in ("valu_busy", "valu_instructions_per_evaluation", "cycles_per_valu_instruction",
              "hardware_tflops"):
        assert c[k] > 0
    assert bench.measured_traffic("integral_persistent_mixed_batch64") > 0


def test_bench_refuses_more_than_one_gpu():
    """One GPU per run: --gpus 2 exits with status 2 and says why, before
    touching a device."""
    import subprocess
    import sys
    from pathlib import Path
    repo = Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, str(repo / "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_bench_rows_share_partitions_every_region():
    """The 8 ranks' shares tile each region's mu rows with the same cells."""
    import bench
    p, _ = synth.workload(1, synth.FULL_SIZE, 3)
    shares = [bench.rows_share(p, r, 8) for r in range(8)]
    for gi, g in enumerate(p.regions):
        step = (g.mu_max - g.mu_min) / g.mu_steps
        parts = [s.regions[gi] for s in shares]
        assert sum(x.mu_steps for x in parts) == g.mu_steps
        assert parts[0].mu_min == g.mu_min and parts[-1].mu_max == pytest.approx(g.mu_max)
        for a, b in zip(parts, parts[1:]):
            assert a.mu_max == pytest.approx(b.mu_min)
        for x in parts:
            assert (x.mu_max - x.mu_min) / x.mu_steps == pytest.approx(step)
            assert (x.r_steps, x.nu_steps) == (g.r_steps, g.nu_steps)
