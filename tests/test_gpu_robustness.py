"""Failure handling on the GPU path: a non-finite state is reported by the step call
(RBN_ENUMERIC), and a failed autotune leaves every ensemble member where it started.

The non-finite state is planted from the host through rbn_set_populations; every kernel
still runs to completion on it (NaN positions of particles are rejected by the range check
before any index is formed)."""
import numpy as np
import pytest

from lagrangianrbnudging_amd import rbn
from scenarios import assert_equal_arrays, conduction_T, random_particles, rb_params, synthetic_observations

pytestmark = pytest.mark.gpu


def _solver(prec, variant, particles):
    p = rb_params(48, 24, 1e5, noutconfig=10)
    p.precision = prec
    if variant == rbn.RBN_VARIANT_NUDGING:
        p.variant = variant
        p.ntempread = 4
        p.alpha = 0.1
    s = rbn.Solver(p)
    s.init_fields(T=conduction_T(p, 1e-2, seed=5))
    if variant == rbn.RBN_VARIANT_NUDGING:
        ids, x0, y0, T, recs = synthetic_observations(48, 24, 30, 200, 4, seed=6)
        s.init_observers(ids, x0, y0, T)
        s.set_observation_table(recs)
    elif particles:
        ids, x, y, tau = random_particles(p, 20, 10, seed=7)
        s.init_particles(ids, x, y, tau_p=tau)
    return s


def _plant(s, value, node=(17, 9), k=3, which="f"):
    s.step(3)
    f, g = s.populations()
    fl = s.fields()
    (f if which == "f" else g)[node[0], node[1], k] = value
    s.set_populations(f, g, fl["rho"], fl["T"], fl["u"], fl["v"])


@pytest.mark.parametrize("prec", [rbn.RBN_FP64, rbn.RBN_FP32])
@pytest.mark.parametrize("variant,particles", [(rbn.RBN_VARIANT_TRAJECTORIES, False),
                                               (rbn.RBN_VARIANT_TRAJECTORIES, True),
                                               (rbn.RBN_VARIANT_NUDGING, False)])
@pytest.mark.parametrize("value,which,nsteps", [(np.nan, "f", 1), (np.inf, "g", 1), (np.nan, "g", 6),
                                                (-np.inf, "f", 5)])
def test_a_non_finite_population_makes_the_step_fail(prec, variant, particles, value, which, nsteps):
    # nsteps 5 and 6 go through two-step launches in the truth run (one step per launch in
    # the nudged run); the planted value reaches the moments of the node that pulls it
    s = _solver(prec, variant, particles)
    _plant(s, value, which=which)
    with pytest.raises(rbn.RbnError) as ei:
        s.step(nsteps)
    assert ei.value.code == rbn.STATUS_CODE["RBN_ENUMERIC"]
    assert "non-finite" in str(ei.value)
    # the state stays non-finite: the next call reports it again
    with pytest.raises(rbn.RbnError) as ei:
        s.step(2)
    assert ei.value.code == rbn.STATUS_CODE["RBN_ENUMERIC"]
    s.close()


@pytest.mark.parametrize("prec", [rbn.RBN_FP64, rbn.RBN_FP32])
def test_finite_runs_report_nothing(prec):
    for variant, particles in ((rbn.RBN_VARIANT_TRAJECTORIES, True), (rbn.RBN_VARIANT_NUDGING, False)):
        s = _solver(prec, variant, particles)
        for n in (1, 2, 7, 1, 12):
            s.step(n)
        assert np.isfinite(s.fields()["T"]).all()
        s.close()


def test_an_ensemble_member_with_a_non_finite_state_fails_the_ensemble_step():
    members = [_solver(rbn.RBN_FP32, rbn.RBN_VARIANT_TRAJECTORIES, True),
               _solver(rbn.RBN_FP32, rbn.RBN_VARIANT_NUDGING, False)]
    _plant(members[1], np.nan, which="g")
    members[0].step(3)
    e = rbn.Ensemble(members)
    e.set_tuning(256, 1, 0, 0)
    with pytest.raises(rbn.RbnError) as ei:
        e.step(4)
    assert ei.value.code == rbn.STATUS_CODE["RBN_ENUMERIC"] and "member 1" in str(ei.value)
    e.close()


def _nudged_members(last_record):
    """Three FP32 nudged members whose observation tables end at step `last_record`."""
    out = []
    for k, (nx, ny, ntr) in enumerate(((48, 24, 5), (64, 32, 5), (40, 24, 3))):
        p = rb_params(nx, ny, 2e5, noutconfig=ntr)
        p.precision = rbn.RBN_FP32
        p.variant = rbn.RBN_VARIANT_NUDGING
        p.ntempread = ntr
        p.alpha = 0.1
        p.deposit = rbn.RBN_DEPOSIT_BILINEAR
        s = rbn.Solver(p)
        s.init_fields(T=conduction_T(p, 1e-2, seed=11 + k))
        ids, x0, y0, T, recs = synthetic_observations(nx, ny, 40, last_record, ntr, seed=21 + k)
        s.init_observers(ids, x0, y0, T)
        s.set_observation_table({st: r for st, r in recs.items() if st <= last_record})
        out.append(s)
    return out


def _state(s):
    f, g = s.populations()
    out = {"f": f, "g": g, **s.fields(), **s.nudging()}
    out["istep"] = np.array(s.istep)
    return out


def test_failed_autotune_restores_every_member():
    # Records reach step 40 only: the autotuner's comparison of the two forms needs at least
    # 2 lcm(6, 5, 3) + 32 = 92 steps and fails on its first missing record.
    members = _nudged_members(40)
    for s in members:
        s.step(2)
    before = [_state(s) for s in members]
    e = rbn.Ensemble(members)
    e.set_tuning(128, 2, 2, 1)
    with pytest.raises(rbn.RbnError) as ei:
        e.autotune(4)
    assert ei.value.code == rbn.STATUS_CODE["RBN_ECALLBACK"]
    assert e.tuning() == {"tile_w": 128, "tile_h": 2, "wg_per_cu": 2, "order": 1}
    for k, (s, b) in enumerate(zip(members, before)):
        a = _state(s)
        for key in b:
            assert_equal_arrays(a[key], b[key], f"member {k} {key}")
    # and the members step on from that state as if the autotuner had never run
    ref = _nudged_members(40)
    for s in ref:
        s.step(2 + 9)
    e.step(9)
    for k, (s, r) in enumerate(zip(members, ref)):
        a, b = _state(s), _state(r)
        for key in b:
            assert_equal_arrays(a[key], b[key], f"member {k} {key} after 9 steps")
    e.close()
