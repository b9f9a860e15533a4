"""The five workload configurations at their own sizes and horizons, through the C-ABI:

C1  256 x 128, FP64, 1000 tracers, records every 10 steps: all 10,000 steps bit for bit
    against the oracle (state, particles and every record).
C2  4096 x 2048, Ra = 1e8, 1e5 tracers + 1e5 inertial particles, records every 20 steps:
    21 FP64 steps bit for bit against the oracle (two-step launches, the record step, the
    carried particle updates); the FP32 shifted-population run against the FP64 run over
    2000 steps within the stated error.
C3  1024 x 512, Ra = 1e7, nudged: N_p in {1e3, 1e4, 1e5} x alpha in {0.01, 0.1, 1}, both
    deposits, FP64 bit for bit against the oracle across record intervals.
C4  the benchmark's 64 FP32 nudged members (sizes from {256x128 ... 1024x512}, Ra
    log-uniform in [1e6, 1e8], 1e3-1e4 tracers each) advanced as one ensemble, with graph
    replays and with eager overlapped steps, against each member stepped on its own; FP64
    members of an ensemble against the oracle directly.
C5  512 x 256, Ra = 1e7, 1e4 tracer observations every 10 steps, window 2000, checkpoints
    every 100: the forward is the nudged run bit for bit (FP32 and FP64), the FP32 gradient
    lies within the stated bound of the FP64 gradient, FP64 central differences on 32 random
    directions match to 1e-5 relative, and the gradient costs at most 4 forward windows.

Measured values are printed (pytest -s) beside each bound."""
import time

import numpy as np
import pytest

import bench
from lagrangianrbnudging_amd import rbn
from oracle.oracle import Oracle
from scenarios import assert_equal_arrays, conduction_T, random_particles, rb_params, setup, synthetic_observations

pytestmark = pytest.mark.gpu


def _compare_populations(a, b, what=""):
    fa, ga = a.populations()
    fb, gb = b.populations()
    assert_equal_arrays(fa, fb, what + "f")
    assert_equal_arrays(ga, gb, what + "g")


def _compare_fields(a, b, what=""):
    x, y = a.fields(), b.fields()
    for k in ("rho", "T", "u", "v"):
        assert_equal_arrays(x[k], y[k], what + k)


def _compare_particles(a, b, what=""):
    x, y = a.particles(), b.particles()
    for k in ("id", "x", "y", "u", "v", "xo", "yo", "uo", "vo"):
        assert_equal_arrays(x[k], y[k], what + "particle " + k)


def _compare_nudging(a, b, what=""):
    x, y = a.nudging(), b.nudging()
    for k in ("nudge_vol", "tempreadint", "nud_term", "weight"):
        assert_equal_arrays(x[k], y[k], what + k)


# ------------------------------------------------------------------------------------- C1


def test_c1_full_horizon_fp64_bit_exact():
    p = rb_params(256, 128, 1e6, noutconfig=10)
    T0 = conduction_T(p, 1e-3, seed=1)
    parts = random_particles(p, 1000, 0, seed=2)
    g = rbn.Solver(p)
    g.collect_records()
    setup(g, p, T0, parts)
    o = setup(Oracle(p), p, T0, parts)
    for _ in range(4):  # calls of 2500 steps
        g.step(2500)
        o.step(2500)
    assert g.istep == o.istep == 10_000
    _compare_populations(g, o)
    _compare_fields(g, o)
    _compare_particles(g, o)
    ra, rb = g.records, o.records
    assert [r["step"] for r in ra] == [r["step"] for r in rb] == list(range(10, 10_001, 10))
    for x, y in zip(ra, rb):
        for k in ("id", "x", "y", "u", "v", "xo", "yo", "uo", "vo", "temp"):
            assert_equal_arrays(x[k], y[k], f"record {x['step']} {k}")
    g.close()


# ------------------------------------------------------------------------------------- C2

NX2, NY2 = 4096, 2048


def _c2(prec):
    p = rb_params(NX2, NY2, 1e8, noutconfig=20, nout=10 ** 9)
    p.precision = prec
    return p


def test_c2_fp64_bit_exact_across_a_record_step():
    # 21 steps: two-step launches carrying the previous pair's particle updates, the record
    # of step 20 (the second step of a pair, its moments written), step 21 on its own.
    p = _c2(rbn.RBN_FP64)
    T0 = conduction_T(p, 1e-3, seed=1)
    parts = random_particles(p, 100_000, 100_000, seed=2)
    g = setup(rbn.Solver(p), p, T0, parts)
    g.collect_records()
    o = setup(Oracle(p), p, T0, parts)
    g.step(21)
    o.step(21)
    _compare_populations(g, o)
    _compare_fields(g, o)
    _compare_particles(g, o)
    assert [r["step"] for r in g.records] == [r["step"] for r in o.records] == [20]
    for k in ("id", "x", "y", "u", "v", "xo", "yo", "uo", "vo", "temp"):
        assert_equal_arrays(g.records[0][k], o.records[0][k], "record " + k)
    g.close()


def test_c2_fp32_error_against_fp64():
    # The FP32 run stores shifted populations (f - w rho0, g - w T0) and sums the deviations;
    # over 2000 steps of the benchmark workload it stays within max|dT| <= 1e-4 (plate
    # difference 1) and max|du| <= 1e-2 max|u| + 1e-7 of the FP64 run, which is the oracle's
    # bit for bit; particle positions within 1e-3 lattice units.
    out = {}
    for prec in (rbn.RBN_FP64, rbn.RBN_FP32):
        p = _c2(prec)
        s = setup(rbn.Solver(p), p, conduction_T(p, 1e-3, seed=1), random_particles(p, 100_000, 100_000, seed=2))
        s.step(2000)
        out[prec] = (s.fields(), s.particles())
        s.close()
    (f64, p64), (f32, p32) = out[rbn.RBN_FP64], out[rbn.RBN_FP32]
    dT = float(np.max(np.abs(f32["T"] - f64["T"])))
    umax = float(max(np.max(np.abs(f64["u"])), np.max(np.abs(f64["v"]))))
    du = float(max(np.max(np.abs(f32["u"] - f64["u"])), np.max(np.abs(f32["v"] - f64["v"]))))
    dxw = np.abs(p32["x"] - p64["x"])
    dx = float(max(np.max(np.minimum(dxw, NX2 - dxw)), np.max(np.abs(p32["y"] - p64["y"]))))  # x periodic
    print(f"C2 FP32 vs FP64 after 2000 steps: max|dT| {dT:.3e}, max|du| {du:.3e} (max|u| {umax:.3e}), "
          f"max particle |dx| {dx:.3e}")
    assert dT <= 1e-4
    assert du <= 1e-2 * umax + 1e-7
    assert dx <= 1e-3


# ------------------------------------------------------------------------------------- C3


@pytest.mark.parametrize("deposit", [rbn.RBN_DEPOSIT_BILINEAR, rbn.RBN_DEPOSIT_BOX7])
@pytest.mark.parametrize("nobs", [1_000, 10_000, 100_000])
def test_c3_nudging_grid_fp64_bit_exact(nobs, deposit):
    # every alpha of the configuration on the same observers; 13 steps cross two record
    # intervals of 5 (snap at 5 and 10, a new record read at 11)
    ids, x0, y0, T, recs = synthetic_observations(1024, 512, nobs, 40, 5, seed=21 + nobs % 7)
    for alpha in (0.01, 0.1, 1.0):
        p = rb_params(1024, 512, 1e7, noutconfig=5)
        p.variant = rbn.RBN_VARIANT_NUDGING
        p.ntempread = 5
        p.alpha = alpha
        p.deposit = deposit
        T0 = conduction_T(p, 1e-2, seed=22)
        sims = []
        for make in (rbn.Solver, Oracle):
            s = make(p)
            s.init_fields(T=T0)
            s.init_observers(ids, x0, y0, T)
            s.set_observation_source(recs)
            s.step(13)
            sims.append(s)
        g, o = sims
        what = f"alpha {alpha}: "
        _compare_populations(g, o, what)
        _compare_fields(g, o, what)
        _compare_nudging(g, o, what)
        g.close()


# ------------------------------------------------------------------------------------- C4


def test_c4_benchmark_ensemble_equals_individual_steps():
    # The 64 members bench.py times (the same seeds), advanced by one ensemble: calls of 1
    # and 92 steps with two graph replays of the 30-step cycle (one launch per step, the
    # passes in turn, from call step 31), then calls of 9 and 31 eager steps with each step's
    # deposit built beside the previous lattice launch (order bit 4); against rbn_step of
    # every member.
    calls_graph, calls_eager = (1, 92), (9, 31)
    total = sum(calls_graph) + sum(calls_eager)
    last = total + 2 * bench.ENS_NTEMPREAD
    ens, nodes, nobs = bench.ens_members(rbn, True, last)
    ref, _, _ = bench.ens_members(rbn, True, last)
    assert len(ens) == 64 and nodes == 13_762_560 and nobs == 352_512
    e = rbn.Ensemble(ens)
    e.set_tuning(128, 4, 4, 2)
    for n in calls_graph:
        e.step(n)
    e.set_tuning(128, 4, 4, 2 | 16)
    for n in calls_eager:
        e.step(n)
    for s in ref:
        s.step(total)
    for k, (a, b) in enumerate(zip(ens, ref)):
        assert a.istep == b.istep == total
        _compare_populations(a, b, f"member {k} ")
        _compare_nudging(a, b, f"member {k} ")
    e.close()


def _fp64_members(kind):
    """A truth run with tracers and inertial particles, a nudged run with the 7x7 deposit
    (records every 3 steps) and one with the bilinear deposit (every 5): FP64, on the GPU
    (kind rbn.Solver) or the oracle."""
    out = []
    p = rb_params(64, 32, 2e5, noutconfig=10 ** 6)
    s = kind(p)
    s.init_fields(T=conduction_T(p, 1e-2, seed=41))
    ids, x, y, tau = random_particles(p, 30, 20, seed=42)
    s.init_particles(ids, x, y, tau_p=tau)
    out.append(s)
    for k, (nx, ny, ntr, dep, alpha) in enumerate(((48, 24, 3, rbn.RBN_DEPOSIT_BOX7, 0.3),
                                                   (40, 32, 5, rbn.RBN_DEPOSIT_BILINEAR, 1.0))):
        q = rb_params(nx, ny, 1e5 * (k + 1), noutconfig=ntr)
        q.variant = rbn.RBN_VARIANT_NUDGING
        q.ntempread = ntr
        q.alpha = alpha
        q.deposit = dep
        s = kind(q)
        s.init_fields(T=conduction_T(q, 1e-2, seed=43 + k))
        oids, x0, y0, T, recs = synthetic_observations(nx, ny, 50 + 20 * k, 200, ntr, seed=45 + k)
        s.init_observers(oids, x0, y0, T)
        s.set_observation_source(recs)
        out.append(s)
    return out


@pytest.mark.parametrize("order", [0, 16])
def test_c4_fp64_ensemble_members_equal_the_oracle(order):
    ens = _fp64_members(rbn.Solver)
    orc = _fp64_members(Oracle)
    e = rbn.Ensemble(ens)
    e.set_tuning(256, 1, 4, order)
    for n in (1, 92, 5):  # cycle of lcm(6, 3, 5) = 30 from call step 31: two replays when graphs are on
        e.step(n)
    for o in orc:
        o.step(98)
    for k, (a, b) in enumerate(zip(ens, orc)):
        _compare_populations(a, b, f"member {k} ")
        _compare_fields(a, b, f"member {k} ")
        if k == 0:
            _compare_particles(a, b, "member 0 ")
        else:
            _compare_nudging(a, b, f"member {k} ")
    e.close()


# ------------------------------------------------------------------------------------- C5

NX5, NY5, EVERY5, NOBS5, ALPHA5, W5, K5 = 512, 256, 10, 10_000, 0.1, 2000, 100


@pytest.fixture(scope="module")
def c5():
    """A truth run (FP64, from its own initial temperature) records 1e4 tracers every 10
    steps; the nudged runs start from another temperature and are nudged toward them."""
    p = rb_params(NX5, NY5, 1e7, noutconfig=EVERY5, nout=10 ** 9)
    rng = np.random.default_rng(21)
    lin = np.tile(p.Tdown + (p.Tup - p.Tdown) * np.arange(NY5) / (NY5 - 1), (NX5, 1))
    T_truth = lin + rng.uniform(-1e-2, 1e-2, lin.shape)
    T_first = lin + rng.uniform(-1e-2, 1e-2, lin.shape)
    ids = np.arange(NOBS5, dtype=np.int32)
    x = rng.uniform(0.0, NX5, NOBS5)
    x[x == 0.0] = 0.5
    y = rng.uniform(1.0, NY5, NOBS5)
    t = rbn.Solver(p)
    t.init_fields(T=T_truth)
    t.init_particles(ids, x, y, tau_p=np.zeros(NOBS5))
    T_at = t.sample("T", x, y)
    t.collect_records()
    t.step(W5 + EVERY5)
    recs = {r["step"]: {"id": r["id"], "x": r["x"], "y": r["y"], "temp": r["temp"]} for r in t.records}
    t.close()
    return {"p": p, "recs": recs, "obs": (ids, x, y, T_at), "T0": T_first}


def _c5_solver(c, prec, alpha=ALPHA5):
    q = c["p"].copy()
    q.variant = rbn.RBN_VARIANT_NUDGING
    q.ntempread = EVERY5
    q.alpha = alpha
    q.precision = prec
    s = rbn.Solver(q)
    s.set_observation_table(c["recs"])
    return s


def _c5_J(s, c, T0):
    s.init_fields(T=T0)
    s.init_observers(*c["obs"])
    return s.adj_forward(W5, K5)


@pytest.mark.parametrize("prec", [rbn.RBN_FP32, rbn.RBN_FP64])
def test_c5_forward_is_the_nudged_run(c5, prec):
    a = _c5_solver(c5, prec)
    _c5_J(a, c5, c5["T0"])
    b = _c5_solver(c5, prec)
    b.init_fields(T=c5["T0"])
    b.init_observers(*c5["obs"])
    b.step(W5)
    assert a.istep == b.istep == W5
    _compare_populations(a, b)
    _compare_fields(a, b)
    _compare_nudging(a, b)
    a.close()
    b.close()


def test_c5_gradient_fp32_against_fp64_and_cost(c5):
    res = {}
    for prec in (rbn.RBN_FP64, rbn.RBN_FP32):
        s = _c5_solver(c5, prec)
        _c5_J(s, c5, c5["T0"])  # warm-up: allocations
        s.adj_backward()
        s.init_fields(T=c5["T0"])
        s.init_observers(*c5["obs"])
        s.synchronize()
        t0 = time.perf_counter()
        J = s.adj_forward(W5, K5)
        g, ga = s.adj_backward()
        t_grad = time.perf_counter() - t0
        r = _c5_solver(c5, prec)
        r.init_fields(T=c5["T0"])
        r.init_observers(*c5["obs"])
        r.synchronize()
        t0 = time.perf_counter()
        r.step(W5)
        r.synchronize()
        t_fwd = time.perf_counter() - t0
        res[prec] = (J, g, ga, t_grad / t_fwd)
        s.close()
        r.close()
    (J64, g64, a64, c64), (J32, g32, a32, c32) = res[rbn.RBN_FP64], res[rbn.RBN_FP32]
    gm = float(np.max(np.abs(g32 - g64)) / np.max(np.abs(g64)))
    print(f"C5 FP32 vs FP64: |dJ|/J {abs(J32 - J64) / J64:.2e}, max|dg|/max|g| {gm:.2e}, "
          f"|d dJ/dalpha| rel {abs(a32 - a64) / abs(a64):.2e}; gradient cost / forward: FP32 {c32:.2f}, FP64 {c64:.2f}")
    # FP32 forward with FP64 adjoint arithmetic and reductions (measured 1e-6, 2.5e-5, 1.5e-5)
    assert abs(J32 - J64) <= 1e-5 * J64
    assert gm <= 2e-4
    assert abs(a32 - a64) <= 2e-4 * abs(a64)
    assert c32 <= 4.0 and c64 <= 4.0


def test_c5_gradient_matches_central_differences_fp64(c5):
    s = _c5_solver(c5, rbn.RBN_FP64)
    _c5_J(s, c5, c5["T0"])
    g, ga = s.adj_backward()
    rng = np.random.default_rng(6)
    eps = 1e-5
    worst = 0.0
    for _ in range(32):
        d = rng.standard_normal(g.shape)
        fd = (_c5_J(s, c5, c5["T0"] + eps * d) - _c5_J(s, c5, c5["T0"] - eps * d)) / (2 * eps)
        ad = float(np.sum(g * d))
        worst = max(worst, abs(fd - ad) / abs(ad))
    epsa = 1e-6
    sp, sm = _c5_solver(c5, rbn.RBN_FP64, ALPHA5 + epsa), _c5_solver(c5, rbn.RBN_FP64, ALPHA5 - epsa)
    fda = (_c5_J(sp, c5, c5["T0"]) - _c5_J(sm, c5, c5["T0"])) / (2 * epsa)
    print(f"C5 central differences (eps 1e-5, 32 directions): worst rel {worst:.2e}; "
          f"dJ/dalpha rel {abs(fda - ga) / abs(ga):.2e}")
    assert worst <= 1e-5
    assert abs(fda - ga) <= 1e-5 * abs(ga)
    for x in (s, sp, sm):
        x.close()
