"""Writes golden.json: the digests of every output of the cases in cases.py, run on the CPU
oracle (FP64).  Run from the repository root after building the oracle:

    python tests/golden/make_golden.py
"""
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE))

from cases import CASES, digest  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402


def main() -> None:
    golden = {}
    for name, case in CASES.items():
        out = case(Oracle)
        golden[name] = {k: digest(v) for k, v in sorted(out.items())}
        print(f"{name}: {len(out)} arrays", flush=True)
    (HERE / "golden.json").write_text(json.dumps(golden, indent=1, sort_keys=True) + "\n")


if __name__ == "__main__":
    main()
