"""Parity of librbn's HIP path with the CPU FP64 oracle (through the C-ABI).

FP64: bit-exact (np.array_equal) on populations, fields, particle state and trajectory
records.  Shapes are small enough that the oracle finishes in seconds.
"""
import numpy as np
import pytest

from lagrangianrbnudging_amd import rbn
from oracle.oracle import Oracle
from scenarios import (assert_equal_arrays, conduction_T, max_rel, random_particles, rb_params,
                       setup)

pytestmark = pytest.mark.gpu


def _run_both(p, T0, particles, steps, collect=True, **tune):
    g = rbn.Solver(p)
    g.set_tuning(**tune)
    if collect:
        g.collect_records()
        g.collect_diagnostics()
    setup(g, p, T0, particles)
    o = setup(Oracle(p), p, T0, particles)
    g.step(steps)
    o.step(steps)
    return g, o


def _compare_state(g, o, what=""):
    fg, gg = g.populations()
    fo, go = o.populations()
    assert_equal_arrays(fg, fo, what + "f")
    assert_equal_arrays(gg, go, what + "g")
    a, b = g.fields(), o.fields()
    for k in ("rho", "T", "u", "v"):
        assert_equal_arrays(a[k], b[k], what + k)


def _compare_particles(g, o):
    a, b = g.particles(), o.particles()
    for k in ("id", "x", "y", "u", "v", "xo", "yo", "uo", "vo"):
        assert_equal_arrays(a[k], b[k], "particle " + k)


def _compare_records(g, o, full):
    ra, rb = g.records, o.records
    assert [r["step"] for r in ra] == [r["step"] for r in rb]
    keys = ["x", "y", "u", "v", "xo", "yo", "uo", "vo", "temp"]
    if full:
        keys += ["dtu", "dtv", "dxu", "dxv", "dyu", "dyv", "fu", "fv"]
    for x, y in zip(ra, rb):
        assert_equal_arrays(x["id"], y["id"], f"record {x['step']} id")
        for k in keys:
            assert_equal_arrays(x[k], y[k], f"record {x['step']} {k}")


@pytest.mark.parametrize("fuse,steps", [(1, 1), (0, 1), (1, 2), (1, 3)])
@pytest.mark.parametrize("writefull", [0, 1])
def test_truth_run_fp64_bit_exact_tiny(writefull, fuse, steps):
    p = rb_params(32, 16, 1e5, noutconfig=7, prtcl_writefull=writefull, nout=50)
    T0 = conduction_T(p, 1e-2, seed=11)
    parts = random_particles(p, 24, 24, seed=12)
    g, o = _run_both(p, T0, parts, 300, particle_fuse=fuse, lbm_steps=steps)
    _compare_state(g, o)
    _compare_particles(g, o)
    _compare_records(g, o, writefull)
    assert len(g.records) == 300 // 7
    # evolution.dat: deterministic device reduction vs sequential sum
    dg, do = g.diagnostics, o.diagnostics
    assert [d[0] for d in dg] == [d[0] for d in do]
    for a, b in zip(dg, do):
        assert abs(a[1] - b[1]) <= 1e-12 * abs(b[1]) + 1e-300
        assert abs(a[2] - b[2]) <= 1e-12 * abs(b[2])


@pytest.mark.parametrize("steps", [1, 2, 3])
@pytest.mark.parametrize("writefull", [0, 1])
def test_particle_fuse_fp64_bit_exact(writefull, steps):
    # The particle update riding in the next lattice launch, including the cases that fall
    # back to a launch of its own: the last step of a call (chunks of 13, 2, 50, 35 steps)
    # and a record step followed by a diagnostics step (records every 7, energy every 5:
    # steps 14/15, 49/50, ...).
    p = rb_params(48, 24, 1e5, noutconfig=7, prtcl_writefull=writefull, nout=5)
    T0 = conduction_T(p, 1e-2, seed=21)
    parts = random_particles(p, 40, 40, seed=22)
    g = rbn.Solver(p)
    g.set_tuning(particle_fuse=1, lbm_steps=steps)
    g.collect_records()
    g.collect_diagnostics()
    setup(g, p, T0, parts)
    o = setup(Oracle(p), p, T0, parts)
    for n in (13, 2, 50, 35):
        g.step(n)
    o.step(100)
    _compare_state(g, o)
    _compare_particles(g, o)
    _compare_records(g, o, writefull)
    assert len(g.records) == 100 // 7
    assert [d[0] for d in g.diagnostics] == [d[0] for d in o.diagnostics]


@pytest.mark.parametrize("writefull", [0, 1])
def test_particle_fuse_fp32_matches_separate_launches(writefull):
    p = rb_params(200, 64, 1e6, noutconfig=10, prtcl_writefull=writefull, nout=11)
    p.precision = rbn.RBN_FP32
    T0 = conduction_T(p, 1e-2, seed=23)
    parts = random_particles(p, 300, 300, seed=24)
    runs = []
    for tune in (dict(particle_fuse=0, lbm_steps=1), dict(particle_fuse=1, lbm_steps=1),
                 dict(particle_fuse=1, lbm_steps=1, lbm_xcd=2), dict(particle_fuse=1, lbm_steps=2),
                 dict(particle_fuse=1, lbm_steps=2, lbm2_seg=0, lbm2_tile_h=4),
                 dict(particle_fuse=1, lbm_steps=2, lbm2_seg=16, lbm2_xcd=0), dict(particle_fuse=0, lbm_steps=2),
                 dict(particle_fuse=1, lbm_steps=3), dict(particle_fuse=1, lbm_steps=3, lbm3_seg=7, lbm2_xcd=0),
                 dict(particle_fuse=1, lbm_steps=3, lbm3_nt=1), dict(particle_fuse=0, lbm_steps=3)):
        g = rbn.Solver(p)
        g.set_tuning(**tune)
        g.collect_records()
        setup(g, p, T0, parts)
        for n in (40, 1, 59):
            g.step(n)
        runs.append(g)
    for other in runs[1:]:
        _compare_state_pair(runs[0], other)
        a, b = runs[0].particles(), other.particles()
        for k in ("x", "y", "u", "v", "xo", "yo", "uo", "vo"):
            assert_equal_arrays(a[k], b[k], "particle " + k)
        assert [r["step"] for r in runs[0].records] == [r["step"] for r in other.records]
        for x, y in zip(runs[0].records, other.records):
            for k in x:
                if isinstance(x[k], np.ndarray):
                    assert_equal_arrays(x[k], y[k], f"record {x['step']} {k}")
    for r in runs:
        r.close()


@pytest.mark.parametrize("prec", [rbn.RBN_FP64, rbn.RBN_FP32])
@pytest.mark.parametrize("shape,tile_h,seg,variant", [((200, 44), 4, 0, 0), ((200, 44), 8, 0, 0), ((200, 44), 16, 0, 0),
                                                      ((40, 20), 8, 0, 0), ((128, 16), 16, 0, 0), ((200, 44), 0, 8, 0),
                                                      ((300, 37), 0, 13, 1), ((40, 20), 0, 64, 0), ((520, 16), 0, 32, 1),
                                                      ((257, 24), 0, 8, 0), ((64, 70), 0, 64, 1), ((200, 44), 0, 43, 1)])
def test_two_step_launch_is_two_steps(prec, shape, tile_h, seg, variant):
    # Tiles (seg 0): partial tiles in x and y, a lattice narrower than one tile (the ring
    # wraps onto the tile itself).  Row sweeps: partial strips (down to one column, whose
    # halo columns are its periodic neighbours), a lattice narrower than a strip, segments
    # that do not divide the rows, a single segment.  Energy diagnostics every 7 steps (a
    # diagnostics step never starts a pair), calls of odd and even lengths.
    nx, ny = shape
    p = rb_params(nx, ny, 1e5, particle=0, nout=7)
    p.precision = prec
    T0 = conduction_T(p, 1e-2, seed=27)
    runs = []
    for steps in (1, 2):
        s = rbn.Solver(p)
        s.set_tuning(lbm_steps=steps, lbm2_tile_h=tile_h, lbm2_seg=seg, lbm2_variant=variant)
        s.collect_diagnostics()
        s.init_fields(T=T0)
        for n in (23, 8, 1, 2, 9):
            s.step(n)
        runs.append(s)
    _compare_state_pair(*runs)
    assert runs[0].diagnostics == runs[1].diagnostics
    if prec == rbn.RBN_FP64:
        o = Oracle(p)
        o.init_fields(T=T0)
        o.step(43)
        _compare_state(runs[1], o)
    for s in runs:
        s.close()


@pytest.mark.parametrize("prec", [rbn.RBN_FP64, rbn.RBN_FP32])
@pytest.mark.parametrize("shape,seg,xcd", [((200, 44), 0, 1), ((200, 44), 8, 0), ((40, 20), 4, 1), ((8, 12), 0, 1),
                                           ((252, 30), 11, 1), ((253, 17), 5, 0), ((520, 16), 64, 1),
                                           ((300, 37), 13, 0), ((64, 70), 4096, 1)])
def test_three_step_launch_is_three_steps(prec, shape, seg, xcd):
    # Row sweeps three steps deep: a lattice of 8 columns (the strip's threads wrap around it
    # many times), exactly one strip (252 columns), two uneven
    # strips (253), several strips, segments of 4 rows up to one segment for the whole
    # lattice, segments that do not divide the rows, both workgroup orders (the hardware order
    # with non-temporal stores, the other with cached ones).  Energy diagnostics
    # every 7 steps (a diagnostics step is never among the first two of a launch), calls of
    # lengths that leave launches of three, two and one step.
    nx, ny = shape
    p = rb_params(nx, ny, 1e5, particle=0, nout=7)
    p.precision = prec
    T0 = conduction_T(p, 1e-2, seed=29)
    runs = []
    for steps in (1, 3):
        s = rbn.Solver(p)
        s.set_tuning(lbm_steps=steps, lbm3_seg=seg, lbm2_xcd=xcd, lbm3_nt=1 - xcd)
        s.collect_diagnostics()
        s.init_fields(T=T0)
        for n in (23, 8, 1, 2, 9, 4):
            s.step(n)
        runs.append(s)
    _compare_state_pair(*runs)
    assert runs[0].diagnostics == runs[1].diagnostics
    if prec == rbn.RBN_FP64:
        o = Oracle(p)
        o.init_fields(T=T0)
        o.step(47)
        _compare_state(runs[1], o)
    for s in runs:
        s.close()


def test_three_step_launch_counts():
    # 31 steps, records every 1000: ten three-step launches, the last step alone (it writes the
    # moments).  Records every 5 over 30 steps: ten three-step launches, a record step at any of
    # their steps (5 and 20 second, 10 and 25 first, 15 and 30 third); the launches of steps
    # 13-15 and 28-30 rewrite T, which the carried record updates of steps 10 and 25 read, so
    # those run first on their own (three launches each), and the last three after the call.
    p = rb_params(64, 32, 1e5, noutconfig=1000)
    g = rbn.Solver(p)
    g.set_tuning(lbm_steps=3)
    assert g.tuning()["lbm_steps"] == 3
    setup(g, p, conduction_T(p, 1e-2, seed=25), random_particles(p, 50, 50, seed=26))
    g.kernel_timing(True, every=1)
    g.step(31)
    assert [g.kernel_stats(k)[0] for k in ("lbm", "lbm2", "lbm3", "lbm_fused", "particles")] == [0, 0, 10, 1, 1]
    g.close()
    p = rb_params(64, 32, 1e5, noutconfig=5)
    g = rbn.Solver(p)
    g.set_tuning(lbm_steps=3)
    setup(g, p, conduction_T(p, 1e-2, seed=25), random_particles(p, 50, 50, seed=26))
    g.kernel_timing(True, every=1)
    g.step(30)
    assert [g.kernel_stats(k)[0] for k in ("lbm3", "lbm2", "lbm", "lbm_fused", "particles")] == [10, 0, 0, 0, 9]
    g.close()


def test_sampled_kernel_timing_counts():
    # every = 3 over one call of 31 steps: the first lattice launch carries nothing, the
    # next 30 carry the previous step's particle update, the last step's update runs alone.
    p = rb_params(64, 32, 1e5, noutconfig=1000)
    g = rbn.Solver(p)
    g.set_tuning(lbm_steps=1)
    setup(g, p, conduction_T(p, 1e-2, seed=25), random_particles(p, 50, 50, seed=26))
    g.kernel_timing(True, every=3)
    g.step(31)
    assert g.kernel_stats("lbm")[0] == 1
    assert g.kernel_stats("lbm_fused")[0] == 10
    assert g.kernel_stats("particles")[0] == 1
    n, ms = g.kernel_stats("lbm_fused", reset=True)
    assert ms > 0 and g.kernel_stats("lbm_fused")[0] == 0
    g.kernel_timing(False)
    g.step(5)
    assert g.kernel_stats("lbm_fused")[0] == 0
    with pytest.raises(rbn.RbnError):
        g.kernel_stats("nope")
    g.close()
    # Two steps per launch (the default): steps 1..30 in 15 two-step launches, the last
    # step (it writes the moments) alone, carrying the updates of steps 29 and 30; the last
    # step's update on its own.
    g = rbn.Solver(p)
    setup(g, p, conduction_T(p, 1e-2, seed=25), random_particles(p, 50, 50, seed=26))
    g.kernel_timing(True, every=3)
    g.step(31)
    assert [g.kernel_stats(k)[0] for k in ("lbm", "lbm2", "lbm_fused", "particles")] == [0, 5, 1, 1]
    g.close()


def _compare_state_pair(a, b):
    fa, ga = a.populations()
    fb, gb = b.populations()
    assert_equal_arrays(fa, fb, "f")
    assert_equal_arrays(ga, gb, "g")
    x, y = a.fields(), b.fields()
    for k in ("rho", "T", "u", "v"):
        assert_equal_arrays(x[k], y[k], k)


def test_truth_run_fp64_bit_exact_config1_shape():
    # 256x128, Ra = 1e6, 1000 tracers (the CPU-runnable reference case), 600 steps.
    p = rb_params(256, 128, 1e6, noutconfig=10)
    T0 = conduction_T(p, 1e-3, seed=1)
    parts = random_particles(p, 1000, 0, seed=2)
    g, o = _run_both(p, T0, parts, 600)
    _compare_state(g, o)
    _compare_particles(g, o)
    _compare_records(g, o, 0)


def test_reference_scratch_init_fp64():
    # The reference's tanh front with srand(0)/rand() jitter, Tdown = 0.01, Tup = -0.01.
    p = rbn.default_params(nx=48, ny=24, scratch=1, noutconfig=5, nout=1000)
    g = rbn.Solver(p)
    g.init_fields()
    o = Oracle(p)
    o.init_scratch_reference()
    _compare_state(g, o, "init ")
    ids = np.arange(16, dtype=np.int32)
    x = np.linspace(1.5, 46.5, 16)
    y = np.linspace(1.2, 23.8, 16)
    g.init_particles(ids, x, y)  # reference kind: heavy, taulag = 2150, 0.98 * fluid velocity
    o.init_particles(ids, x, y)
    g.step(50)
    o.step(50)
    _compare_state(g, o)
    _compare_particles(g, o)


def test_restart_round_trip_fp64():
    p = rb_params(40, 20, 2e5, noutconfig=1000)
    T0 = conduction_T(p, 1e-2, seed=3)
    a = rbn.Solver(p)
    a.init_fields(T=T0)
    a.step(120)
    f, gpop = a.populations()
    fl = a.fields()
    a.step(80)
    b = rbn.Solver(p)
    b.set_populations(f, gpop, fl["rho"], fl["T"], fl["u"], fl["v"])
    b.step(80)
    fa, ga = a.populations()
    fb, gb = b.populations()
    assert_equal_arrays(fa, fb, "f")
    assert_equal_arrays(ga, gb, "g")
    # and the oracle restarted from the same state agrees
    o = Oracle(p)
    o.set_populations(f, gpop, fl["rho"], fl["T"], fl["u"], fl["v"])
    o.step(80)
    fo, go = o.populations()
    assert_equal_arrays(fb, fo, "f oracle")
    assert_equal_arrays(gb, go, "g oracle")


def test_no_particles_and_energy():
    p = rb_params(64, 32, 1e5, particle=1)
    T0 = conduction_T(p, 1e-2, seed=5)
    g = rbn.Solver(p)
    g.init_fields(T=T0)
    g.init_particles(np.zeros(0, np.int32), np.zeros(0), np.zeros(0))
    o = Oracle(p)
    o.init_fields(T=T0)
    g.step(100)
    o.step(100)
    _compare_state(g, o)
    eg, tg = g.energy()
    eo, to = o.energy()
    assert abs(eg - eo) <= 1e-12 * eo and abs(tg - to) <= 1e-12 * abs(to)


@pytest.mark.parametrize("prec", [rbn.RBN_FP64, rbn.RBN_FP32])
def test_launch_tuning_does_not_change_results(prec):
    p = rb_params(200, 40, 1e5)  # nx not a multiple of any block width: partial segments
    p.precision = prec
    T0 = conduction_T(p, 1e-2, seed=13)
    ref = None
    shapes = [(b, nt, 0, x, 0) for b in (64, 128, 256, 512) for nt in (0, 1) for x in (0, 1)]
    shapes += [(256, 1, 1, 0, 0), (256, 1, 0, 0, 1), (512, 1, 1, 1, 1), (64, 0, 1, 0, 1)]
    shapes = [s + (1, 0) for s in shapes] + [(256, 1, 0, 0, 1, 2, 0), (128, 1, 0, 1, 1, 4, 0), (64, 1, 0, 0, 0, 8, 0)]
    shapes += [(256, 1, 0, 0, 1, 1, 1), (256, 1, 1, 0, 1, 1, 1), (256, 0, 0, 0, 1, 1, 1)]
    shapes += [(256, 1, 0, 2, 1, 1, 1), (128, 1, 0, 2, 0, 2, 0), (512, 1, 0, 2, 1, 1, 1)]
    for block, nt, ntload, xcd, rev, npt, keep in shapes:
        s = rbn.Solver(p)
        s.set_tuning(lbm_block=block, lbm_nt=nt, lbm_ntload=ntload, lbm_xcd=xcd, lbm_reverse=rev, lbm_npt=npt,
                     lbm_uv_keep=keep)
        s.init_fields(T=T0)
        s.step(30)
        f, g = s.populations()
        fl = s.fields()
        what = f"block={block} nt={nt} ntload={ntload} xcd={xcd} reverse={rev} npt={npt} uv_keep={keep}"
        if ref is None:
            ref = (f, g, fl)
        else:
            assert_equal_arrays(f, ref[0], f"f {what}")
            assert_equal_arrays(g, ref[1], f"g {what}")
            for k in ("u", "v", "T", "rho"):
                assert_equal_arrays(fl[k], ref[2][k], f"{k} {what}")
        s.close()


@pytest.mark.parametrize("prec", [rbn.RBN_FP64, rbn.RBN_FP32])
def test_row_profiles_match_the_oracle(prec):
    p = rb_params(96, 48, 1e6)
    T0 = conduction_T(p, 5e-2, seed=17)
    q = p.copy()
    q.precision = prec
    g = rbn.Solver(q)
    g.init_fields(T=T0)
    o = Oracle(p)
    o.init_fields(T=T0)
    g.step(120)
    o.step(120)
    a, b = g.profiles(), o.profiles()
    assert a.shape == (48, 9)
    err = np.max(np.abs(a - b), axis=0)
    scale = np.max(np.abs(b), axis=0)
    if prec == rbn.RBN_FP64:
        # the same values summed in another order
        assert np.all(err <= 1e-13 * scale + 1e-300), err
    else:
        # FP32 fields against FP64: the stated field error (1e-4 on T, 1e-2 relative on u, v)
        assert np.all(err <= 2e-2 * scale + 2e-4), err
    assert np.max(np.abs(b[:, 8])) > 0  # vorticity developed


def test_two_contexts_are_independent():
    p = rb_params(32, 16, 1e5)
    a, b = rbn.Solver(p), rbn.Solver(p)
    a.init_fields(T=conduction_T(p, 1e-2, 7))
    b.init_fields(T=conduction_T(p, 1e-2, 7))
    a.step(10)
    b.step(5)
    b.step(5)
    assert_equal_arrays(a.populations()[1], b.populations()[1], "g")


def test_particle_leaving_lattice_is_an_error():
    p = rb_params(32, 16, 1e5)
    g = rbn.Solver(p)
    g.init_fields(T=conduction_T(p, 1e-2, 7))
    with pytest.raises(rbn.RbnError) as e:
        g.init_particles(np.array([0], np.int32), np.array([-5.0]), np.array([4.0]))
    assert e.value.code == -6


def test_fp32_shifted_error_against_fp64():
    # FP32 with shifted populations against the FP64 oracle after 2000 steps at 128x64,
    # Ra = 1e6 (convective, growing perturbation; measured max|dT| = 3.7e-5 with dT_plates = 1).
    # Stated bound: max |T32 - T64| <= 1e-4 and max |u32 - u64| <= 1e-2 * max|u64| + 1e-7.
    p64 = rb_params(128, 64, 1e6)
    p32 = p64.copy()
    p32.precision = rbn.RBN_FP32
    T0 = conduction_T(p64, 1e-3, seed=9)
    g = rbn.Solver(p32)
    g.init_fields(T=T0)
    o = Oracle(p64)
    o.init_fields(T=T0)
    g.step(2000)
    o.step(2000)
    a, b = g.fields(), o.fields()
    errT = np.max(np.abs(a["T"] - b["T"]))
    erru = np.max(np.abs(a["u"] - b["u"]))
    print(f"FP32 vs FP64: max|dT| {errT:.3e}, max|du| {erru:.3e} (max|u| {np.max(np.abs(b['u'])):.3e}), "
          f"rel rho {max_rel(a['rho'], b['rho']):.3e}")
    assert errT <= 1e-4
    assert erru <= 1e-2 * np.max(np.abs(b["u"])) + 1e-7
