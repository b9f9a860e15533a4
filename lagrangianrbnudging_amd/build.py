"""Build recipe of librbn (HIP, gfx950) and of the C host driver.

Compiles every source under csrc/ with hipcc for gfx950 into lagrangianrbnudging_amd/librbn.so
(in-tree, so it travels with the repository), and the drop-in driver rbn_driver (C) against it.
Run as ``python -m lagrangianrbnudging_amd.build`` or through ``__graft_entry__.build()``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "librbn.so"
DRIVER = PKG / "rbn_driver"

ARCH = os.environ.get("RBN_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: the FP64 path reproduces the reference's un-fused arithmetic bit for
# bit.  Packed-FP32 instruction selection is disabled for all device code.
HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build librbn)")


def _run_parallel(cmds: list[list[str]]) -> None:
    procs = [(c, subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)) for c in cmds]
    errors = []
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            errors.append(f"$ {' '.join(cmd)}\n{out}")
        elif out and "error" in out.lower():
            print(out, file=sys.stderr)
    if errors:
        raise RuntimeError("build failed:\n" + "\n".join(errors))


def build_library(verbose: bool = False) -> Path:
    hipcc = _hipcc()
    OBJ.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    cmds, objs = [], []
    for src in sources:
        obj = OBJ / (src.name + ".o")
        lang = ["-x", "hip"] if src.suffix == ".hip" else []
        cmds.append([hipcc, *HIP_FLAGS, *lang, "-c", str(src), "-o", str(obj)])
        objs.append(str(obj))
    if verbose:
        for c in cmds:
            print(" ".join(c))
    _run_parallel(cmds)
    tmp = LIB.with_name(LIB.name + ".tmp")
    link = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs]
    _run_parallel([link])
    os.replace(tmp, LIB)
    return LIB


def build_driver() -> Path:
    """The C host program that replays the reference's main() call sequence over librbn."""
    src = PKG / "driver" / "rbn_driver.c"
    if not src.exists():
        return DRIVER
    cc = shutil.which("gcc") or "cc"
    cmd = [cc, "-std=c99", "-O2", "-Wall", "-o", str(DRIVER), str(src),
           f"-I{ROOT / 'include'}", f"-L{PKG}", "-lrbn", "-ldl", "-lm", f"-Wl,-rpath,$ORIGIN"]
    _run_parallel([cmd])
    return DRIVER


def build_oracle() -> Path:
    """The CPU restatement used by the tests as the checker (never linked into librbn)."""
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    return ROOT / "oracle" / "_build" / "liboracle.so"


def build_all(verbose: bool = False) -> None:
    build_library(verbose)
    build_driver()
    build_oracle()


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print(f"built {LIB}")
