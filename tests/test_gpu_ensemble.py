"""Batched ensemble: persistent launches over all members (one per step carrying the previous
step's particle update, a lattice and a particle launch per step, or one per two steps on
two-step tiles carrying the previous launch's two particle updates), hipGraph-captured, must
give exactly what stepping every member on its own gives."""
import numpy as np
import pytest

from lagrangianrbnudging_amd import rbn
from scenarios import assert_equal_arrays, conduction_T, random_particles, rb_params, synthetic_observations

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
                                    (256, 2, 8, 10), (128, 4, 2, 15), (256, 1, 0, 16), (128, 4, 4, 20), (256, 1, 4, 26)])
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


@pytest.mark.parametrize("prec", [rbn.RBN_FP64, rbn.RBN_FP32])
@pytest.mark.parametrize("order", [0, 8])
def test_ensemble_after_members_step_on_their_own(prec, order):
    # Members stepped on their own between ensemble calls (three steps per launch, which
    # leaves a member's latest velocity past the ensemble's six-slot ring), all of them or one
    # of them: the next ensemble call sees the changed members (rebuilds its descriptors and
    # graph) and continues exactly as stepping every member on its own does.
    ens_members = _members(prec)
    ref_members = _members(prec)
    for s in ens_members + ref_members:
        s.set_tuning(lbm_steps=3)
    e = rbn.Ensemble(ens_members)
    e.set_tuning(256, 1, 0, order)
    for n, own in ((12, None), (7, "all"), (12, None), (4, 3), (24, None), (5, "all"), (13, None)):
        if own is None:
            e.step(n)
            for s in ref_members:
                s.step(n)
        elif own == "all":
            for a, b in zip(ens_members, ref_members):
                a.step(n)
                b.step(n)
        else:
            ens_members[own].step(n)
            ref_members[own].step(n)
    _compare(ens_members, ref_members)
    e.close()


# Nudged members: (nx, ny, Ra, observers, ntempread, deposit, alpha)
NUDGED = [(48, 24, 1e5, 40, 4, rbn.RBN_DEPOSIT_BOX7, 0.3), (64, 40, 3e5, 90, 5, rbn.RBN_DEPOSIT_BILINEAR, 1.0),
          (40, 24, 5e4, 0, 5, rbn.RBN_DEPOSIT_BILINEAR, 0.1), (72, 32, 2e5, 25, 3, rbn.RBN_DEPOSIT_BOX7, 0.05)]


def _mixed_members(prec, table, intervals=None):
    """Two truth-run members with particles, then the nudged members (one without observers);
    observations through the library's table (ensemble) or a Python provider (reference).
    intervals: the nudged members' record intervals instead of NUDGED's."""
    out = []
    for k, (nx, ny, ra) in enumerate(SHAPES[:2]):
        p = rb_params(nx, ny, ra, noutconfig=10 ** 6)
        p.precision = prec
        s = rbn.Solver(p)
        s.init_fields(T=conduction_T(p, 1e-2, seed=60 + k))
        ids, x, y, tau = random_particles(p, 30, 15, seed=70 + k)
        s.init_particles(ids, x, y, tau_p=tau)
        out.append(s)
    for k, (nx, ny, ra, nobs, ntr, dep, alpha) in enumerate(NUDGED):
        if intervals is not None:
            ntr = intervals[k]
        p = rb_params(nx, ny, ra, noutconfig=ntr)
        p.precision = prec
        p.variant = rbn.RBN_VARIANT_NUDGING
        p.ntempread = ntr
        p.alpha = alpha
        p.deposit = dep
        s = rbn.Solver(p)
        s.init_fields(T=conduction_T(p, 1e-2, seed=80 + k))
        if nobs > 0:
            ids, x0, y0, T, recs = synthetic_observations(nx, ny, nobs, 420, ntr, seed=90 + k)
            s.init_observers(ids, x0, y0, T)
        else:  # no observers: records with no entries
            none = {"id": np.zeros(0, np.int32), "x": np.zeros(0), "y": np.zeros(0), "temp": np.zeros(0)}
            s.init_observers(none["id"], none["x"], none["y"], none["temp"])
            recs = {st: none for st in range(ntr, 420 + ntr, ntr)}
        if table:
            s.set_observation_table(recs)
        else:
            s.set_observation_source(recs)
        out.append(s)
    return out


def _compare_nudged(a, b, first_nudged=2):
    for k, (sa, sb) in enumerate(zip(a, b)):
        fa, ga = sa.populations()
        fb, gb = sb.populations()
        assert_equal_arrays(fa, fb, f"member {k} f")
        assert_equal_arrays(ga, gb, f"member {k} g")
        for key in ("rho", "T", "u", "v"):
            assert_equal_arrays(sa.fields()[key], sb.fields()[key], f"member {k} {key}")
        assert sa.istep == sb.istep
        if k >= first_nudged:  # the deposit of the last step and the nudging term
            na, nb = sa.nudging(), sb.nudging()
            for key in na:
                assert_equal_arrays(na[key], nb[key], f"member {k} {key}")
        else:
            pa, pb = sa.particles(), sb.particles()
            for key in ("x", "y", "u", "v", "xo", "yo", "uo", "vo"):
                assert_equal_arrays(pa[key], pb[key], f"member {k} particle {key}")


@pytest.mark.parametrize("prec", [rbn.RBN_FP64, rbn.RBN_FP32])
@pytest.mark.parametrize("tuning,intervals", [((256, 1, 0, 0), None), ((128, 4, 4, 3), None), ((64, 8, 2, 1), None),
                                              ((256, 1, 4, 10), None), ((256, 1, 0, 2), (3, 3, 2, 6)),
                                              ((128, 4, 4, 1), (3, 3, 2, 6)), ((256, 1, 0, 18), (3, 3, 2, 6)),
                                              ((256, 1, 0, 8), (3, 3, 2, 6)), ((256, 1, 2, 9), None),
                                              ((256, 1, 4, 27), (3, 3, 2, 6)), ((256, 1, 0, 4), None)])
def test_ensemble_with_nudged_members_equals_individual_steps(prec, tuning, intervals):
    # Nudged members (7x7 and bilinear deposits, different record intervals, one without
    # observers) beside truth-run members; the observer updates and deposits of all nudged
    # members run as batched launches, records are uploaded between steps or, inside the
    # captured cycles, scattered by the graph from a staging set uploaded before each replay.
    # One launch per step: cycles of 60 steps for intervals 4, 5, 5, 3, of 6 for 3, 3, 2, 6,
    # from call step 31.  Two steps per launch (order bit 3), both steps' deposits before it:
    # cycles of 60 and 12 from call step 32.  Calls of 122 steps (cycles and eager steps), 120
    # (whole cycles: the graphs stay for the next call) and 93 (a replay of the kept graph).
    # Separate-launch orders (bit 2) fall back to one launch per step.
    ens_members = _mixed_members(prec, table=True, intervals=intervals)
    ref_members = _mixed_members(prec, table=False, intervals=intervals)
    e = rbn.Ensemble(ens_members)
    e.set_tuning(*tuning)
    calls = (9, 1, 2, 6, 13, 20, 122, 120, 93)
    for n in calls:
        e.step(n)
    for s in ref_members:
        s.step(sum(calls))
    _compare_nudged(ens_members, ref_members)
    e.close()


@pytest.mark.parametrize("prec", [rbn.RBN_FP64, rbn.RBN_FP32])
@pytest.mark.parametrize("order", [0, 8, 10, 24])
def test_ensemble_of_nudged_members_only_equals_individual_steps(prec, order):
    # An ensemble of nudged members only runs the two-step kernel compiled for nudged tiles
    # alone (order bit 3); odd calls end on a one-step launch and leave the last deposit in
    # the second set.  The call of 46 steps replays the captured cycle (6 or 12 steps from
    # call step 31 or 32) where graphs are on.
    ens_members = _mixed_members(prec, table=True, intervals=(3, 3, 2, 6))[2:]
    ref_members = _mixed_members(prec, table=False, intervals=(3, 3, 2, 6))[2:]
    e = rbn.Ensemble(ens_members)
    e.set_tuning(256, 1, 4, order)
    calls = (1, 2, 15, 12, 46, 27)
    for n in calls:
        e.step(n)
    for s in ref_members:
        s.step(sum(calls))
    _compare_nudged(ens_members, ref_members, first_nudged=0)
    e.close()


def test_ensemble_autotune_with_nudged_members_keeps_state():
    ens_members = _mixed_members(rbn.RBN_FP32, table=True)
    ref_members = _mixed_members(rbn.RBN_FP32, table=False)
    e = rbn.Ensemble(ens_members)
    assert e.autotune(4) > 0
    # one launch per step or two (bit 3); not the separate-launch orders (bit 2); bit 4 (eager
    # steps) or graph replays, whichever measured faster
    assert (e.tuning()["order"] & 4) == 0
    for s in ens_members:
        assert s.istep == 0
    e.step(17)
    for s in ref_members:
        s.step(17)
    _compare_nudged(ens_members, ref_members)
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
    with pytest.raises(rbn.RbnError):
        rbn.Ensemble([a, a])
    n = p.copy()
    n.variant = rbn.RBN_VARIANT_NUDGING
    n.ntempread = 4
    c = rbn.Solver(n)
    c.init_fields(T=conduction_T(n, 1e-2, 1))
    ids, x0, y0, T, _ = synthetic_observations(32, 16, 10, 20, 4, seed=3)
    c.init_observers(ids, x0, y0, T)
    e = rbn.Ensemble([a, c])  # a nudged member without an observation source fails at its first record
    with pytest.raises(rbn.RbnError):
        e.step(3)
    e.close()
