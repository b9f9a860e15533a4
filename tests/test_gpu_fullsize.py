"""The benchmark-scale configuration itself (4096 x 2048, Ra = 1e8, 1e5 tracers + 1e5 inertial
particles): a few FP64 steps bit for bit against the oracle (it finishes them in seconds),
and size-independent properties over longer FP32 runs — two-step launches equal one-step
launches bit for bit, mass is conserved, particles stay inside the lattice and finite."""
import numpy as np
import pytest

from lagrangianrbnudging_amd import rbn
from oracle.oracle import Oracle
from scenarios import assert_equal_arrays, conduction_T, random_particles, rb_params, setup

pytestmark = pytest.mark.gpu

NX, NY = 4096, 2048


def _params(prec, noutconfig):
    p = rb_params(NX, NY, 1e8, noutconfig=noutconfig, nout=10 ** 9)
    p.precision = prec
    return p


def _particles_inside(sim):
    q = sim.particles()
    for k in ("x", "y", "u", "v", "xo", "yo", "uo", "vo"):
        assert np.isfinite(q[k]).all(), k
    assert (q["x"] >= 0.0).all() and (q["x"] < NX).all()
    assert (q["y"] >= 1.0).all() and (q["y"] <= NY).all()


def test_fullsize_fp64_bit_exact_against_oracle():
    # 3 steps with a record at step 2: a two-step launch carrying nothing (steps 1-2, the
    # second a record step), then one step carrying both particle updates, then the last
    # step's update on its own.
    p = _params(rbn.RBN_FP64, 2)
    T0 = conduction_T(p, 1e-3, seed=1)
    parts = random_particles(p, 100_000, 100_000, seed=2)
    g = setup(rbn.Solver(p), p, T0, parts)
    g.collect_records()
    o = setup(Oracle(p), p, T0, parts)
    g.step(3)
    o.step(3)
    fg, gg = g.populations()
    fo, go = o.populations()
    assert_equal_arrays(fg, fo, "f")
    assert_equal_arrays(gg, go, "g")
    del fg, gg, fo, go
    a, b = g.fields(), o.fields()
    for k in ("rho", "T", "u", "v"):
        assert_equal_arrays(a[k], b[k], k)
    pa, pb = g.particles(), o.particles()
    for k in ("id", "x", "y", "u", "v", "xo", "yo", "uo", "vo"):
        assert_equal_arrays(pa[k], pb[k], "particle " + k)
    assert [r["step"] for r in g.records] == [r["step"] for r in o.records] == [2]
    for k in ("id", "x", "y", "u", "v", "temp"):
        assert_equal_arrays(g.records[0][k], o.records[0][k], "record " + k)
    # mass: rho = 1 at the start, conserved by collision, streaming and the walls
    assert abs(a["rho"].sum() / (NX * NY) - 1.0) < 1e-12
    g.close()


def test_fullsize_fp32_two_step_equals_one_step_and_conserves_mass():
    # one, two and three steps per launch (three: the default at this size) with the
    # benchmark's particles and records every 20 steps, calls of 17 and 24 steps
    p = _params(rbn.RBN_FP32, 20)
    T0 = conduction_T(p, 1e-3, seed=3)
    parts = random_particles(p, 100_000, 100_000, seed=4)
    runs = []
    for steps in (1, 2, 3):
        s = rbn.Solver(p)
        s.set_tuning(lbm_steps=steps)
        s.collect_records()
        setup(s, p, T0, parts)
        for n in (17, 24):
            s.step(n)
        runs.append(s)
    d = rbn.Solver(p)
    assert d.tuning()["lbm_steps"] == 3
    d.close()
    fa, ga = runs[0].populations()
    for other in runs[1:]:
        fb, gb = other.populations()
        assert_equal_arrays(fa, fb, "f")
        assert_equal_arrays(ga, gb, "g")
        del fb, gb
        pa, pb = runs[0].particles(), other.particles()
        for k in ("x", "y", "u", "v", "xo", "yo", "uo", "vo"):
            assert_equal_arrays(pa[k], pb[k], "particle " + k)
        assert [r["step"] for r in runs[0].records] == [r["step"] for r in other.records] == [20, 40]
        for x, y in zip(runs[0].records, other.records):
            for k in ("x", "y", "u", "v", "temp"):
                assert_equal_arrays(x[k], y[k], f"record {x['step']} {k}")
    del fa, ga
    rho = runs[2].fields()["rho"]
    assert abs(rho.sum() / (NX * NY) - 1.0) < 1e-6
    _particles_inside(runs[2])
    for s in runs:
        s.close()
