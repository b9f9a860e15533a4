"""The nudging and ensemble configurations at their own sizes: a nudged FP64 run at
1024 x 512 with 1e4 observers bit for bit against the oracle (both deposits), and an
ensemble of members drawn from {256x128 ... 1024x512} (nudged and truth-run, FP32) against
the same members stepped one by one."""
import numpy as np
import pytest

from lagrangianrbnudging_amd import rbn
from oracle.oracle import Oracle
from scenarios import assert_equal_arrays, conduction_T, random_particles, rb_params, synthetic_observations

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("deposit", [rbn.RBN_DEPOSIT_BILINEAR, rbn.RBN_DEPOSIT_BOX7])
def test_nudging_config_fp64_bit_exact(deposit):
    # 1e4 observers on 1024 x 512: the bilinear deposit folds over its active-node list,
    # the 7x7 one every node in row order; 12 steps cross two record intervals of 5.
    p = rb_params(1024, 512, 1e7, noutconfig=5)
    p.variant = rbn.RBN_VARIANT_NUDGING
    p.ntempread = 5
    p.alpha = 0.1
    p.deposit = deposit
    ids, x0, y0, T, recs = synthetic_observations(p.nx, p.ny, 10_000, 40, 5, seed=21)
    T0 = conduction_T(p, 1e-2, seed=22)
    sims = []
    for make in (rbn.Solver, Oracle):
        s = make(p)
        s.init_fields(T=T0)
        s.init_observers(ids, x0, y0, T)
        s.set_observation_source(recs)
        s.step(12)
        sims.append(s)
    g, o = sims
    fg, gg = g.populations()
    fo, go = o.populations()
    assert_equal_arrays(fg, fo, "f")
    assert_equal_arrays(gg, go, "g")
    a, b = g.nudging(), o.nudging()
    for k in ("nudge_vol", "tempreadint", "nud_term"):
        assert_equal_arrays(a[k], b[k], k)
    a, b = g.fields(), o.fields()
    for k in ("rho", "T", "u", "v"):
        assert_equal_arrays(a[k], b[k], k)


SIZES = [(256, 128), (512, 256), (768, 384), (1024, 512)]


def _config_members(table):
    rng = np.random.default_rng(5)
    out = []
    for m in range(16):
        nx, ny = SIZES[(m // 2) % 4]  # every size nudged and truth-run
        ra = 10.0 ** rng.uniform(6.0, 8.0)
        n = int(rng.integers(1000, 10001))
        nudged = m % 2 == 0
        p = rb_params(nx, ny, ra, noutconfig=5 if nudged else 10 ** 6, nout=10 ** 9)
        p.precision = rbn.RBN_FP32
        if nudged:
            p.variant = rbn.RBN_VARIANT_NUDGING
            p.ntempread = 5
            p.alpha = 0.1
            p.deposit = rbn.RBN_DEPOSIT_BILINEAR
        s = rbn.Solver(p)
        s.init_fields(T=conduction_T(p, 1e-3, seed=100 + m))
        if nudged:
            # records up to step 160: the autotuner's longest span (2 * 60 + 33 steps with two
            # steps per launch) reads the record of step 155
            ids, x0, y0, T, recs = synthetic_observations(nx, ny, n, 160, 5, seed=200 + m)
            s.init_observers(ids, x0, y0, T)
            if table:
                s.set_observation_table(recs)
            else:
                s.set_observation_source(recs)
        else:
            pid, x, y, tau = random_particles(p, n, 0, seed=300 + m)
            s.init_particles(pid, x, y, tau_p=tau)
        out.append(s)
    return out


def test_ensemble_config_sizes_equal_individual_steps():
    ens = _config_members(table=True)
    ref = _config_members(table=False)
    e = rbn.Ensemble(ens)
    e.autotune(4)
    # 93: eager steps, a replay of the captured cycle (30 steps from call step 31 with one
    # launch per step, 60 from call step 32 with two) and eager steps
    for n in (1, 31, 93):
        e.step(n)
    for s in ref:
        s.step(125)
    for k, (a, b) in enumerate(zip(ens, ref)):
        fa, ga = a.populations()
        fb, gb = b.populations()
        assert_equal_arrays(fa, fb, f"member {k} f")
        assert_equal_arrays(ga, gb, f"member {k} g")
        if k % 2 == 0:
            na, nb = a.nudging(), b.nudging()
            for key in na:
                assert_equal_arrays(na[key], nb[key], f"member {k} {key}")
        else:
            pa, pb = a.particles(), b.particles()
            for key in ("x", "y", "u", "v"):
                assert_equal_arrays(pa[key], pb[key], f"member {k} particle {key}")
    e.close()


@pytest.mark.parametrize("prec", [rbn.RBN_FP64, rbn.RBN_FP32])
def test_deposit_overlap_equals_in_turn(prec):
    # dep_overlap 1 (deposits on a second stream, two deposit sets by step parity) against 0,
    # over calls of odd and even lengths, both ending on either set; the getters too.
    p = rb_params(256, 128, 1e7, noutconfig=5)
    p.precision = prec
    p.variant = rbn.RBN_VARIANT_NUDGING
    p.ntempread = 5
    p.alpha = 0.1
    p.deposit = rbn.RBN_DEPOSIT_BILINEAR
    ids, x0, y0, T, recs = synthetic_observations(p.nx, p.ny, 2000, 40, 5, seed=31)
    T0 = conduction_T(p, 1e-2, seed=32)
    sims = []
    for ovl in (0, 1):
        s = rbn.Solver(p)
        s.set_tuning(dep_overlap=ovl)
        s.init_fields(T=T0)
        s.init_observers(ids, x0, y0, T)
        s.set_observation_table(recs)
        for n in (1, 2, 7, 12, 5, 3):
            s.step(n)
            sims.append((ovl, n, s.nudging()))
        fa, ga = s.populations()
        sims.append((ovl, "final", {"f": fa, "g": ga}))
        s.close()
    half = len(sims) // 2
    for (_, n, a), (_, _, b) in zip(sims[:half], sims[half:]):
        for k in a:
            assert_equal_arrays(a[k], b[k], f"after call {n}: {k}")
