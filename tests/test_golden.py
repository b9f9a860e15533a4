"""The committed golden fixtures (tests/golden/golden.json, written by make_golden.py from the
oracle): the oracle must still reproduce them (CPU), and the HIP path must reproduce them bit
for bit in FP64 (GPU)."""
import json
import sys
from pathlib import Path

import pytest

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))

from cases import CASES, digest  # noqa: E402
from lagrangianrbnudging_amd import rbn  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

GOLDEN = json.loads((Path(__file__).resolve().parent / "golden" / "golden.json").read_text())


def _check(out, name):
    want = GOLDEN[name]
    assert sorted(out) == sorted(want), f"{name}: outputs {sorted(set(out) ^ set(want))}"
    bad = []
    for k, v in out.items():
        d = digest(v)
        if d["shape"] != want[k]["shape"] or d["sha256"] != want[k]["sha256"]:
            bad.append(f"{k}: shape {d['shape']} vs {want[k]['shape']}, first {d['first']} vs {want[k]['first']}, "
                       f"sum {d['sum']!r} vs {want[k]['sum']!r}")
    assert not bad, f"{name}: " + "; ".join(bad)


def test_golden_file_covers_every_case():
    assert sorted(GOLDEN) == sorted(CASES)


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_golden(name):
    _check(CASES[name](Oracle), name)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_hip_path_reproduces_golden(name):
    sims = []

    def make(p):
        s = rbn.Solver(p)
        sims.append(s)
        return s

    try:
        _check(CASES[name](make), name)
    finally:
        for s in sims:
            s.close()
