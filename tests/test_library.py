"""The C-ABI library loads, exports every declared entry point, and refuses to run without a GPU."""
import ctypes
import re
from pathlib import Path

import pytest

from lagrangianrbnudging_amd import rbn

ROOT = Path(__file__).resolve().parents[1]


def header_functions():
    text = (ROOT / "include" / "rbn.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(rbn_[a-z0-9_]+)\s*\(", text))
    typedef_fns = set(re.findall(r"\(\*\s*(rbn_[a-z0-9_]+)\)", text))
    return sorted(names - typedef_fns)


def test_every_declared_function_is_exported():
    lib = ctypes.CDLL(str(rbn.LIB_PATH))
    declared = header_functions()
    assert len(declared) >= 30
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    assert sorted(rbn.declared_symbols()) == header_functions()


def test_version_and_api():
    lib = rbn.load()
    assert lib.rbn_api_version() == 1
    assert b"gfx950" in lib.rbn_version()


def test_params_struct_matches_c_layout():
    # The ctypes mirror must agree with the C struct: round-trip every field through
    # rbn_params_set and read it back.
    p = rbn.default_params()
    lib = rbn.load()
    for name, value in [("nx", 77), ("Tup", -0.25), ("heavy_init_factor", 0.5)]:
        pass
    assert lib.rbn_params_set(ctypes.byref(p), b"nx", 77.0) == 0
    assert lib.rbn_params_set(ctypes.byref(p), b"rbn_heavy_init_factor", 0.5) == 0
    assert lib.rbn_params_set(ctypes.byref(p), b"Tup", -0.25) == 0
    assert lib.rbn_params_set(ctypes.byref(p), b"rbn_ndiag", 12.0) == 0
    assert (p.nx, p.Tup, p.heavy_init_factor, p.ndiag) == (77, -0.25, 0.5, 12)


def test_create_without_gpu_fails_loudly():
    if rbn.device_count() > 0:
        pytest.skip("a GPU is visible")
    p = rbn.default_params(nx=32, ny=16)
    with pytest.raises(rbn.RbnError) as e:
        rbn.Solver(p)
    assert e.value.code == -4
    assert "no GPU" in str(e.value)


def test_invalid_params_rejected_before_device():
    p = rbn.default_params(nx=4, ny=16)
    with pytest.raises(rbn.RbnError) as e:
        rbn.Solver(p)
    assert e.value.code == -1
    p = rbn.default_params(nx=32, ny=16, relax=0.4)
    with pytest.raises(rbn.RbnError):
        rbn.Solver(p)
    p = rbn.default_params(nx=32, ny=16, variant=1, ntempread=1)
    with pytest.raises(rbn.RbnError):
        rbn.Solver(p)
    p = rbn.default_params(nx=32, ny=16, variant=1, ntempread=5, nstart=3)
    with pytest.raises(rbn.RbnError) as e:
        rbn.Solver(p)
    assert "multiple of ntempread" in str(e.value)
