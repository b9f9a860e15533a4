"""Host-side logic of bench.py (no GPU): byte accounting, the record-step count, the
multi-GPU refusal and the slab-decomposed CPU baseline on a small lattice."""
import subprocess
import sys
from pathlib import Path

import pytest

import bench
from lagrangianrbnudging_amd import rbn

ROOT = Path(__file__).resolve().parents[1]


def test_record_steps_counts_multiples_in_the_window():
    assert bench.record_steps(100, 2000) == 100
    assert bench.record_steps(19, 2) == 1      # steps 20, 21
    assert bench.record_steps(0, 19) == 0
    assert bench.record_steps(50, 1000) == 50  # 60, 80, ..., 1040


def test_launch_bytes_are_the_lattice_and_particle_figures():
    nodes = bench.NX * bench.NY
    assert bench.lbm_bytes_per_node(1) == 160 and bench.lbm_bytes_per_node(0) == 320
    # tracer 16*8 + 8 reals, inertial 16*8 + 48 reals
    assert bench.particle_bytes_per_step(1, 1, 0) == 128 + 32
    assert bench.particle_bytes_per_step(1, 0, 1) == 128 + 192
    assert bench.particle_bytes_per_step(0, 1, 1) == 128 + 64 + 128 + 384
    assert bench.launch_bytes(1, 0.0, False) == nodes * 160
    assert bench.launch_bytes(1, 0.05, True) == nodes * (160 + 0.05 * 4) + bench.particle_bytes_per_step(1)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr and r.stdout == ""


def test_slab_cpu_baseline_places_every_particle(monkeypatch):
    monkeypatch.setattr(bench, "NX", 128)
    monkeypatch.setattr(bench, "NY", 64)
    monkeypatch.setattr(bench, "N_TRACER", 300)
    monkeypatch.setattr(bench, "N_INERTIAL", 200)
    r = bench.run_cpu_baseline(rbn, 2, 4)
    assert r["cores"] == 4 and r["kind"] == "port" and r["particles"] == 500
    assert r["value"] > 0 and "4 threads" in r["sample"]


def _autotune_records(trial, ntempread):
    """The records rbn_ensemble_autotune reads from the starting state, restated from its
    loop: every candidate runs a warm-up call of G steps and a timed call of reps * G steps
    (reps = max(1, ceil(trial / G)); G = 6 one launch per step, 12 two steps per launch) and
    is restored; then per form a call of 2 L + 32 steps (L = lcm(6, ntempread)) or 2 L + 33
    (L = lcm(12, ntempread)), whichever launch mode won.  A call reads at step s
    (s % ntempread == 1) the record of step s + ntempread - 1."""
    import math

    spans = []
    for g, t0 in ((6, 31), (12, 32)):
        reps = max(1, (trial + g - 1) // g)
        L = g * ntempread // math.gcd(g, ntempread)
        spans += [g + reps * g, 2 * L + t0 + 1]
    return {s + ntempread - 1 for span in spans for s in range(1, span + 1) if s % ntempread == 1}


@pytest.mark.parametrize("ntempread", [2, 3, 5, 7, 10, 12])
@pytest.mark.parametrize("trial", [1, 6, 8, 13, 40])
def test_ensemble_record_table_covers_the_autotuner(trial, ntempread):
    # the driver's short run (--steps 20 --warmup 5) once asked for record 65 of a table
    # that ended at step 58
    need = max(_autotune_records(trial, ntempread))
    assert bench.ens_autotune_steps(trial, ntempread) + ntempread - 1 >= need
    horizon = bench.ens_record_horizon(trial, 20, 24, 8, ntempread)
    assert horizon >= need
    import numpy as np

    _, _, _, _, recs = bench.ens_observations(np.random.default_rng(0), 64, 32, 4, horizon, ntempread)
    assert all(r in recs for r in _autotune_records(trial, ntempread))
    # the timed run itself: ens_warm + steps and the sequential warmup + steps
    assert max(recs) >= 24 + 8 + ntempread - 1 and max(recs) >= 20 + 8 + ntempread - 1


def test_a_failing_leg_is_reported_inside_the_line(capsys):
    def boom():
        raise RuntimeError("observation provider failed for record 65")

    r = bench.run_leg("demo", boom)
    assert r == {"error": "RuntimeError: observation provider failed for record 65"}
    assert not bench.ok(r) and bench.ok({"mlups": 1.0})
    assert "demo failed" in capsys.readouterr().err
    assert bench.lattice_summary(r, "lbm2_avg_ms", 1.0) is r


def test_launch_minimum_bytes_never_exceed_the_algorithmic_bytes():
    # two one-step launches move 2 * 40 reals per node; a two-step launch at least 42, a
    # three-step one 44 (every population and the force read and written once, the velocity
    # of each step written)
    for prec, s in ((0, 8), (1, 4)):
        nodes = bench.NX * bench.NY
        assert bench.launch_min_bytes(prec, 0.0, False) == nodes * 42 * s
        assert bench.launch_min_bytes(prec, 0.0, False, steps=3) == nodes * 44 * s
        assert bench.launch_min_bytes(prec, 0.05, True) < bench.launch_bytes(prec, 0.05, True, steps=2)
        assert bench.launch_min_bytes(prec, 0.0, True, steps=3) < bench.launch_bytes(prec, 0.0, True, steps=3)
