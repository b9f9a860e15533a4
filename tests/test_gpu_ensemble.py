"""Batched ensemble: persistent launches over all members (one per step carrying the previous
step's particle update, a lattice and a particle launch per step, or one per two steps on
two-step tiles carrying the previous launch's two particle updates), hipGraph-captured, must
give exactly what stepping every member on its own gives."""
import numpy as np
import pytest

from lagrangianrbnudging_amd import rbn
from scenarios import assert_equal_arrays, conduction_T, random_particles, rb_params

pytestmark = pytest.mark.gpu

SHAPES = [(64, 32, 1e5), (96, 48, 3e5), (40, 24, 5e4), (128, 64, 1e6), (72, 36, 2e5)]


def _members(prec, with_particles=True):
    out = []
    for k, (nx, ny, ra) in enumerate(SHAPES):
        p = rb_params(nx, ny, ra, noutconfig=10 ** 6)
        p.precision = prec
        s = rbn.Solver(p)
        s.init_fields(T=conduction_T(p, 1e-2, seed=30 + k))
        if with_particles and k != 2:  # one member without particles
            ids, x, y, tau = random_particles(p, 20 + 7 * k, 10 + 3 * k, seed=40 + k)
            s.init_particles(ids, x, y, tau_p=tau)
        out.append(s)
    return out


def _compare(a, b):
    for k, (sa, sb) in enumerate(zip(a, b)):
        fa, ga = sa.populations()
        fb, gb = sb.populations()
        assert_equal_arrays(fa, fb, f"member {k} f")
        assert_equal_arrays(ga, gb, f"member {k} g")
        for key in ("rho", "T", "u", "v"):
            assert_equal_arrays(sa.fields()[key], sb.fields()[key], f"member {k} {key}")
        pa, pb = sa.particles(), sb.particles()
        for key in ("x", "y", "u", "v", "xo", "yo", "uo", "vo"):
            assert_equal_arrays(pa[key], pb[key], f"member {k} particle {key}")
        assert sa.istep == sb.istep


@pytest.mark.parametrize("prec", [rbn.RBN_FP64, rbn.RBN_FP32])
@pytest.mark.parametrize("tuning", [(256, 1, 0, 0), (256, 2, 4, 1), (64, 8, 2, 2), (128, 4, 8, 3), (256, 1, 0, 2),
                                    (256, 1, 4, 4), (128, 4, 2, 7), (256, 1, 0, 8), (128, 4, 4, 11), (64, 8, 2, 9),
                                    (256, 2, 8, 10), (128, 4, 2, 15)])
def test_ensemble_equals_individual_steps(prec, tuning):
    ens_members = _members(prec)
    ref_members = _members(prec)
    e = rbn.Ensemble(ens_members)
    e.set_tuning(*tuning)
    # Calls of 9, 1, 2, 6, 13, 27, 12, 15 steps: graph cycles (one replay covers six steps
    # with one step per launch, twelve with two) and eager launches, calls without a cycle,
    # parity changes (rebuilds) between calls, a call (12 steps) after which the captured
    # graph is replayed again by the next call.
    calls = (9, 1, 2, 6, 13, 27, 12, 15)
    for n in calls:
        e.step(n)
    for s in ref_members:
        s.step(sum(calls))
    _compare(ens_members, ref_members)
    e.close()


def test_ensemble_autotune_keeps_member_state():
    ens_members = _members(rbn.RBN_FP64)
    ref_members = _members(rbn.RBN_FP64)
    e = rbn.Ensemble(ens_members)
    ms = e.autotune(4)
    assert ms > 0
    t = e.tuning()
    assert t["tile_w"] in (64, 128, 256) and (t["tile_w"] * t["tile_h"]) % 256 == 0
    for s in ens_members:
        assert s.istep == 0
    e.step(8)
    for s in ref_members:
        s.step(8)
    _compare(ens_members, ref_members)


def test_ensemble_rejects_unsupported_members():
    p = rb_params(32, 16, 1e5)
    a = rbn.Solver(p)
    a.init_fields(T=conduction_T(p, 1e-2, 1))
    q = p.copy()
    q.precision = rbn.RBN_FP32
    b = rbn.Solver(q)
    b.init_fields(T=conduction_T(q, 1e-2, 1))
    with pytest.raises(rbn.RbnError):
        rbn.Ensemble([a, b])
    n = p.copy()
    n.variant = rbn.RBN_VARIANT_NUDGING
    n.ntempread = 4
    c = rbn.Solver(n)
    c.init_fields(T=conduction_T(n, 1e-2, 1))
    with pytest.raises(rbn.RbnError):
        rbn.Ensemble([a, c])
    with pytest.raises(rbn.RbnError):
        rbn.Ensemble([a, a])
