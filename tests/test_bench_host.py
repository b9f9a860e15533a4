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
