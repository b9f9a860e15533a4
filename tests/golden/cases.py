"""Tiny seeded cases whose FP64 outputs are committed as golden fixtures (golden.json).

Each case builds its inputs from seeds, runs them on a simulator made by `make(params)` (the
CPU oracle or librbn's Solver: both expose the same methods) and returns named arrays in the
reference's layouts ([nx][ny], [nx][ny][9], particles in id order, records concatenated in
step order).  make_golden.py writes the oracle's outputs; the tests compare the oracle (CPU)
and the HIP path (GPU) against them: the FP64 step is bit-exact, so every array's SHA-256
must match.
"""
from __future__ import annotations

import hashlib

import numpy as np

from lagrangianrbnudging_amd import rbn
from scenarios import conduction_T, random_particles, rb_params, synthetic_observations

RECORD_KEYS = ("id", "x", "y", "u", "v", "xo", "yo", "uo", "vo", "temp")
PARTICLE_KEYS = ("id", "x", "y", "u", "v", "xo", "yo", "uo", "vo")
FULL_KEYS = ("dtu", "dtv", "dxu", "dxv", "dyu", "dyv", "fu", "fv")


def _state(sim, out):
    f, g = sim.populations()
    out["f"], out["g"] = f, g
    for k, v in sim.fields().items():
        out[k] = v


def _records(sim, out, full):
    recs = sim.records
    out["record_steps"] = np.array([r["step"] for r in recs], dtype=np.int32)
    for k in RECORD_KEYS + (FULL_KEYS if full else ()):
        out["record_" + k] = np.concatenate([np.asarray(r[k]) for r in recs])


def truth_writefull(make):
    """32x16, Ra = 1e5, 24 tracers + 24 inertial particles, records with the Writefull fields
    every 7 steps, 200 steps."""
    p = rb_params(32, 16, 1e5, noutconfig=7, prtcl_writefull=1, nout=50)
    sim = make(p)
    if hasattr(sim, "collect_records"):
        sim.collect_records()
    sim.init_fields(T=conduction_T(p, 1e-2, seed=11))
    ids, x, y, tau = random_particles(p, 24, 24, seed=12)
    sim.init_particles(ids, x, y, tau_p=tau)
    sim.step(200)
    out = {}
    _state(sim, out)
    parts = sim.particles()
    for k in PARTICLE_KEYS:
        out["particle_" + k] = parts[k]
    _records(sim, out, True)
    return out


def truth_reference_scratch(make):
    """48x24 with the reference's scratch init (tanh front, glibc srand(0)/rand() jitter) and
    16 particles of the reference kind (heavy, 0.98 of the fluid velocity), 50 steps."""
    p = rbn.default_params(nx=48, ny=24, scratch=1, noutconfig=5, nout=1000)
    sim = make(p)
    if hasattr(sim, "init_scratch_reference"):
        sim.init_scratch_reference()
    else:
        sim.init_fields()
    out = {"T_init": sim.fields()["T"]}
    ids = np.arange(16, dtype=np.int32)
    sim.init_particles(ids, np.linspace(1.5, 46.5, 16), np.linspace(1.2, 23.8, 16))
    sim.step(50)
    _state(sim, out)
    parts = sim.particles()
    for k in PARTICLE_KEYS:
        out["particle_" + k] = parts[k]
    return out


def _nudged(make, deposit, alpha):
    p = rb_params(32, 16, 1e5, noutconfig=4)
    p.variant = rbn.RBN_VARIANT_NUDGING
    p.ntempread = 4
    p.alpha = alpha
    p.deposit = deposit
    ids, x0, y0, T, recs = synthetic_observations(32, 16, 40, 100, 4, seed=31)
    sim = make(p)
    sim.init_fields(T=conduction_T(p, 1e-2, seed=32))
    sim.init_observers(ids, x0, y0, T)
    sim.set_observation_source(recs)
    sim.step(61)
    out = {}
    _state(sim, out)
    for k, v in sim.nudging().items():
        out["nudging_" + k] = v
    return out


def nudged_box7(make):
    """32x16 nudged run, 40 synthetic observers (crossing the periodic seam and the walls),
    records every 4 steps, the reference's 7x7 deposit, alpha = 0.3, 61 steps."""
    return _nudged(make, rbn.RBN_DEPOSIT_BOX7, 0.3)


def nudged_bilinear(make):
    """As nudged_box7 with the bilinear deposit and alpha = 1 (saturated weights)."""
    return _nudged(make, rbn.RBN_DEPOSIT_BILINEAR, 1.0)


CASES = {f.__name__: f for f in (truth_writefull, truth_reference_scratch, nudged_box7, nudged_bilinear)}


def digest(a) -> dict:
    """SHA-256 of the array's little-endian bytes (FP64 for reals, int32 for integers) with
    its shape and a few values for reading a mismatch."""
    a = np.asarray(a)
    a = np.ascontiguousarray(a, dtype="<i4" if a.dtype.kind in "iu" else "<f8")
    flat = a.ravel()
    return {"shape": list(a.shape), "sha256": hashlib.sha256(a.tobytes()).hexdigest(),
            "first": flat[:3].tolist(), "sum": float(flat.astype(np.float64).sum())}
