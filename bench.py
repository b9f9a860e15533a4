#!/usr/bin/env python3
"""Benchmark of the Rayleigh–Bénard step on one MI355X.

Workload (the benchmark-scale plain configuration): 4096 x 2048 lattice, Ra = 1e8, Pr = 1,
linear conduction profile + seeded uniform noise (1e-3) on T, u = 0; 1e5 tracers + 1e5
inertial particles (Stokes number log-uniform in [0.01, 1], tau_p = St * t_ff with
t_ff = (ny-1)/u_ff, zero initial slip), uniform random positions; trajectory records every
20 steps, written to device memory (the rate with every record also copied to the host over
PCIe is reported beside it).  One "step" = the fused f/g LBM update + the particle update
(+ the record on record steps); by default one lattice launch advances three steps (a row
sweep passing the first two steps' populations through LDS; two steps where a record step
ends a launch) and carries the previous launch's particle updates as extra workgroups.
Inputs are resident on the GPU before timing starts.

Prints one JSON line: MLUPS of the FP32 (shifted-population) path, the HBM roofline of the
dominant kernel (the three-step lattice launch; HIP events around launches in the timed
region; the HBM traffic per launch comes from the committed rocprofv3 counter summary under
profiles/), the same workload at two steps per launch (64 x 8 tiles), at one step per
launch and with the particle update as a launch of its own (the one-step lattice kernel's
roofline alone),
the FP64 path's MLUPS on the same workload, the batched ensemble (64 uneven members with
1e3-1e4 nudged tracers each; also truth-run members advecting their tracers) against the
same members stepped one by one, the nudging configuration (1024 x 512: the L2
temperature error of nudged runs against the truth over 20,000 steps, and a sweep over the
number of observed tracers and alpha), the adjoint window's cost against the plain nudged
forward (512 x 256, 2000 steps, checkpoint every 100), and the CPU FP64 restatement of the
reference loop timed on a bounded sample on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "fwd MLUPS & %HBM peak; ensemble MLUPS + adjoint/fwd cost vs plain, 1 GPU"
HBM_PEAK_GBPS = 8000.0
NX, NY = 4096, 2048
RA, PR, U_FF = 1e8, 1.0, 0.05
N_TRACER, N_INERTIAL = 100_000, 100_000
RECORD_EVERY = 20


def make_params(rbn, precision):
    p = rbn.default_params(nx=NX, ny=NY, Tup=0.0, Tdown=1.0, scratch=1, nstart=0, nout=10**9,
                           noutconfig=RECORD_EVERY, prtcl_writefull=0, particle=1,
                           precision=precision, T0_shift=0.5, rho0_shift=1.0)
    return rbn.params_from_ra(p, RA, PR, U_FF)


def make_inputs(p, nx=NX, ny=NY, n_tracer=N_TRACER, n_inertial=N_INERTIAL):
    rng = np.random.default_rng(1)
    j = np.arange(ny)
    T0 = np.tile(p.Tdown + (p.Tup - p.Tdown) * j / (ny - 1), (nx, 1)) + rng.uniform(-1e-3, 1e-3, (nx, ny))
    n = n_tracer + n_inertial
    r3 = np.random.default_rng(3)
    x = r3.uniform(0.0, nx, n)
    x[x == 0.0] = 0.5
    y = r3.uniform(1.0, ny, n)
    r4 = np.random.default_rng(4)
    t_ff = (ny - 1) / U_FF
    st = 10.0 ** r4.uniform(-2.0, 0.0, n_inertial)
    tau = np.concatenate([np.zeros(n_tracer), st * t_ff])
    ids = np.arange(n, dtype=np.int32)
    zeros = np.zeros(n)
    return T0, (ids, x, y, tau, zeros)


def lbm_bytes_per_node(precision):
    # 18 population loads + 18 stores, lagged force load + store, u and v stores.
    return 40 * (8 if precision == 0 else 4)


def particle_bytes_per_step(precision, n_tracer=N_TRACER, n_inertial=N_INERTIAL):
    """Algorithmic bytes of one particle update (record writes excluded): FP64 state, 8 values
    read and 8 written (128 B), plus the fluid values of the interpolation stencil, 8 reals
    for a tracer ((u, v) at the 4 cell corners) and 48 for an inertial particle (the stencil
    of u, v and their time and space differences).  The update itself reads fewer: 5 state
    values (xo, yo, uo, vo are only written), and for an inertial particle (u, v) at 12 nodes
    plus the previous step's at the 4 corners (32 reals)."""
    r = 8 if precision == 0 else 4
    return n_tracer * (16 * 8 + 8 * r) + n_inertial * (16 * 8 + 48 * r)


def record_steps(warmup, steps):
    """Record steps (istep % RECORD_EVERY == 0) among the timed steps warmup+1 .. warmup+steps."""
    return (warmup + steps) // RECORD_EVERY - warmup // RECORD_EVERY


# Every TIMING_EVERY-th launch of each kernel class is bracketed with HIP events in the timed
# region (events between two launches widen the gap between them by a few microseconds).
# 7 is coprime with the record interval, so the sampled launches include record steps in
# proportion.
TIMING_EVERY = 7


def run_gpu(rbn, precision, steps, warmup, timing=True, records_to_host=False, **tune):
    """Time `steps` steps of the benchmark workload.  The particle update writes every record
    into a device slot; with records_to_host each one is also copied to pinned host memory
    (asynchronously, beside later steps) and handed to a callback that drops it."""
    p = make_params(rbn, precision)
    T0, (ids, x, y, tau, zeros) = make_inputs(p)
    s = rbn.Solver(p, 0)
    s.set_tuning(**tune)
    s.init_fields(T=T0)
    s.init_particles(ids, x, y, tau_p=tau, up=zeros, vp=zeros)
    if records_to_host:
        s.discard_records()
    s.step(warmup)
    s.synchronize()
    if timing:
        s.kernel_timing(True, every=TIMING_EVERY)
        for k in ("lbm", "lbm_fused", "lbm2", "lbm3", "particles"):
            s.kernel_stats(k, reset=True)
    t0 = time.perf_counter()
    s.step(steps)
    s.synchronize()
    wall = time.perf_counter() - t0
    out = {"wall_s": wall, "steps": steps, "mlups": NX * NY * steps / wall / 1e6, "ms_per_step": 1e3 * wall / steps}
    if timing:
        # sampled launch counts and summed durations per kernel class
        n_lbm, ms_lbm = s.kernel_stats("lbm")
        n_f, ms_f = s.kernel_stats("lbm_fused")
        n_2, ms_2 = s.kernel_stats("lbm2")
        n_3, ms_3 = s.kernel_stats("lbm3")
        n_p, ms_p = s.kernel_stats("particles")
        s.kernel_timing(False)
        out.update(lbm_timed=n_lbm, lbm_avg_ms=ms_lbm / max(n_lbm, 1),
                   fused_timed=n_f, fused_avg_ms=ms_f / max(n_f, 1),
                   lbm2_timed=n_2, lbm2_avg_ms=ms_2 / max(n_2, 1),
                   lbm3_timed=n_3, lbm3_avg_ms=ms_3 / max(n_3, 1),
                   particles_timed=n_p, particles_avg_ms=ms_p / max(n_p, 1),
                   record_fraction=record_steps(warmup, steps) / steps)
    fl = s.fields()
    out["finite"] = bool(np.isfinite(fl["T"]).all() and np.isfinite(fl["u"]).all())
    s.close()
    return out


ENS_SIZES = [(256, 128), (512, 256), (768, 384), (1024, 512)]
ENS_MEMBERS = 64
ENS_NTEMPREAD, ENS_ALPHA = 5, 0.1


def ens_observations(rng, nx, ny, n, last_step, every):
    """Synthetic observation records of n drifting tracers (straight lines, periodic in x,
    folded into [1, ny] in y, temperatures drifting slowly), one every `every` steps up to
    last_step: the initial state and a step -> record table."""
    x0 = rng.uniform(0.0, nx, n)
    y0 = rng.uniform(1.0, ny, n)
    vx = rng.uniform(-0.05, 0.05, n)
    vy = rng.uniform(-0.03, 0.03, n)
    T0 = rng.uniform(0.0, 1.0, n)
    ids = np.arange(n, dtype=np.int32)
    recs = {}
    for st in range(every, last_step + every, every):
        recs[st] = {"id": ids, "x": np.mod(x0 + vx * st, nx), "y": 1.0 + np.mod(y0 + vy * st - 1.0, ny - 1.0),
                    "temp": np.clip(T0 + 1e-4 * st, 0.0, 1.0)}
    return ids, x0, y0, T0, recs


def ens_members(rbn, nudged, last_step, seed=7):
    """64 FP32 members, sizes drawn from ENS_SIZES, Ra log-uniform in [1e6, 1e8], own seeds,
    1e3..1e4 tracers each: nudged members whose observers follow synthetic records (the
    bilinear deposit, alpha 0.1, records every 5 steps), or truth runs advecting the tracers."""
    rng = np.random.default_rng(seed)
    members, nodes, nobs = [], 0, 0
    for m in range(ENS_MEMBERS):
        nx, ny = ENS_SIZES[int(rng.integers(len(ENS_SIZES)))]
        ra = 10.0 ** rng.uniform(6.0, 8.0)
        n = int(rng.integers(1000, 10001))
        p = rbn.default_params(nx=nx, ny=ny, Tup=0.0, Tdown=1.0, scratch=1, nstart=0, nout=10**9,
                               noutconfig=10**9, prtcl_writefull=0, particle=0 if nudged else 1,
                               precision=rbn.RBN_FP32, T0_shift=0.5, rho0_shift=1.0)
        if nudged:
            p.variant = rbn.RBN_VARIANT_NUDGING
            p.ntempread = ENS_NTEMPREAD
            p.alpha = ENS_ALPHA
            p.deposit = rbn.RBN_DEPOSIT_BILINEAR
            p.noutconfig = ENS_NTEMPREAD
        p = rbn.params_from_ra(p, ra, PR, U_FF)
        s = rbn.Solver(p, 0)
        s.init_conduction(1e-3, 100 + m)
        if nudged:
            ids, x0, y0, T0, recs = ens_observations(rng, nx, ny, n, last_step, ENS_NTEMPREAD)
            s.init_observers(ids, x0, y0, T0)
            s.set_observation_table(recs)
        else:
            x = rng.uniform(0.0, nx, n)
            x[x == 0.0] = 0.5
            y = rng.uniform(1.0, ny, n)
            s.init_particles(np.arange(n, dtype=np.int32), x, y, tau_p=np.zeros(n))
        members.append(s)
        nodes += nx * ny
        nobs += n
    return members, nodes, nobs


ENS_TRIAL_STEPS = 8
# steps of a call before its first graph replay with nudged members (one launch per step;
# one more with two): eager steps during which the first replay's records are fetched
ENS_REPLAY_LEAD = 31


def ens_autotune_steps(trial, ntempread):
    """Steps from its starting state over which rbn_ensemble_autotune advances nudged members
    whose records come every `ntempread` steps (include/rbn.h): each candidate's warm-up call
    of G steps and timed call of G * max(1, ceil(trial / G)) steps (G = 6 for one launch per
    step, 12 for two steps per launch), and the comparison of the two forms over 2 L + 32
    steps, L = lcm(6, ntempread), or 2 L + 33 steps, L = lcm(12, ntempread); the state is
    restored after each."""
    spans = []
    for g, t0 in ((6, ENS_REPLAY_LEAD), (12, ENS_REPLAY_LEAD + 1)):
        lcm = g * ntempread // int(np.gcd(g, ntempread))
        spans += [g + g * max(1, -(-trial // g)), 2 * lcm + t0 + 1]
    return max(spans)


def ens_record_horizon(trial, warmup, ens_warm, steps, ntempread):
    """The last step whose observation record a nudged ensemble leg reads: the sequential
    reference (warmup + steps), the autotuner (restores the state afterwards) and the ensemble
    run (ens_warm + steps), each reading at step s with s % ntempread == 1 the record of step
    s + ntempread - 1."""
    advanced = max(warmup + steps, ens_autotune_steps(trial, ntempread), ens_warm + steps)
    return advanced + ntempread - 1


def run_ensemble(rbn, steps, warmup, tuning=None, nudged=True):
    """The batched ensemble (autotuned, or the given launch shape (tile_w, tile_h, wg_per_cu,
    order)) against the same members, built afresh, stepped one after another with their
    own launches (rbn_step)."""
    # a whole number of the captured cycles (L = lcm(12, ntempread) steps covers both launch
    # modes), long enough to capture them (the lead and a cycle), so that the timed call
    # replays graphs kept from the warm-up call
    cyc = 12 * ENS_NTEMPREAD // int(np.gcd(12, ENS_NTEMPREAD))
    ens_warm = -(-max(warmup, ENS_REPLAY_LEAD + 2 + cyc) // cyc) * cyc
    last = ens_record_horizon(ENS_TRIAL_STEPS, warmup, ens_warm, steps, ENS_NTEMPREAD)
    members, nodes, nobs = ens_members(rbn, nudged, last)
    for s in members:
        s.step(warmup)
    for s in members:
        s.synchronize()
    t0 = time.perf_counter()
    for s in members:
        s.step(steps)
    for s in members:
        s.synchronize()
    t_seq = time.perf_counter() - t0
    for s in members:
        s.close()
    members, nodes, nobs = ens_members(rbn, nudged, last)
    e = rbn.Ensemble(members)
    if tuning is None:
        best_trial = e.autotune(ENS_TRIAL_STEPS)
    else:
        e.set_tuning(*tuning)
        best_trial = float("nan")
    e.step(ens_warm)
    t0 = time.perf_counter()
    e.step(steps)
    t_ens = time.perf_counter() - t0
    out = {"members": ENS_MEMBERS, "nudged": nudged, "nodes": nodes,
           ("observers" if nudged else "tracers"): nobs, "steps": steps,
           "mlups": round(nodes * steps / t_ens / 1e6, 1),
           "sequential_mlups": round(nodes * steps / t_seq / 1e6, 1),
           "speedup_vs_sequential": round(t_seq / t_ens, 2),
           "ms_per_step": round(1e3 * t_ens / steps, 4),
           "roofline_frac": round(nodes * lbm_bytes_per_node(1) / (t_ens / steps) / 1e9 / HBM_PEAK_GBPS, 4),
           "tuning": e.tuning(), "autotune_best_ms": finite_or_null(round(best_trial, 4))}
    order = out["tuning"]["order"]
    out["two_steps_per_launch"] = bool(order & 8) and not order & 4
    out["graph_replays"] = not order & 16
    e.close()
    for s in members:
        s.close()
    return out


NUD_NX, NUD_NY, NUD_RA = 1024, 512, 1e7
NUD_OBS, NUD_EVERY, NUD_ALPHA = 10_000, 5, 0.1
NUD_SWEEP_OBS, NUD_SWEEP_ALPHA = (1_000, 10_000, 100_000), (0.01, 0.1, 1.0)


def nudging_truth(rbn, spinup):
    """The spun-up truth run of the nudging configuration and the nudged runs' first state."""
    pt = rbn.default_params(nx=NUD_NX, ny=NUD_NY, Tup=0.0, Tdown=1.0, scratch=1, nstart=0, nout=10**9,
                            noutconfig=NUD_EVERY, prtcl_writefull=0, particle=1,
                            precision=rbn.RBN_FP32, T0_shift=0.5, rho0_shift=1.0)
    pt = rbn.params_from_ra(pt, NUD_RA, PR, U_FF)
    lin = np.tile(pt.Tdown + (pt.Tup - pt.Tdown) * np.arange(NUD_NY) / (NUD_NY - 1), (NUD_NX, 1))
    # four convection rolls seeded in the truth so that the spin-up reaches developed flow
    ii, jj = np.meshgrid(np.arange(NUD_NX), np.arange(NUD_NY), indexing="ij")
    rolls = 0.05 * np.sin(np.pi * jj / (NUD_NY - 1)) * np.cos(2 * np.pi * 2 * ii / NUD_NX)
    T_truth0 = lin + rolls + np.random.default_rng(5).uniform(-1e-3, 1e-3, lin.shape)
    T_first = lin + np.random.default_rng(6).uniform(-1e-3, 1e-3, lin.shape)
    truth = rbn.Solver(pt, 0)
    truth.init_fields(T=T_truth0)
    truth.step(spinup)
    return pt, truth, T_first


def nudged_solver(rbn, pt, recs, ids, x, y, T_at, T_first, prec, alpha, deposit):
    q = pt.copy()
    q.variant = rbn.RBN_VARIANT_NUDGING
    q.ntempread = NUD_EVERY
    q.alpha = alpha
    q.precision = prec
    q.deposit = deposit
    s = rbn.Solver(q, 0)
    s.set_observation_table(recs)  # held on the device from the first step
    s.init_fields(T=T_first)
    s.init_observers(ids, x, y, T_at)
    return s


def run_nudging_sweep(rbn, steps, spinup):
    """The nudging configuration's parameter sweep: N_p in {1e3, 1e4, 1e5} tracer observations
    times alpha in {0.01, 0.1, 1}, FP32, bilinear deposit, `steps` steps from the same spun-up
    truth.  The observers of the smaller sets are the first ids of one 1e5-tracer truth run
    (tracers do not act on the flow).  Reported: the final L2 temperature error of every run,
    the initial error and the nudged step time per N_p at alpha = 0.1."""
    pt, truth, T_first = nudging_truth(rbn, spinup)
    T_start = truth.fields()["T"]
    n_max = NUD_SWEEP_OBS[-1]
    rng = np.random.default_rng(8)
    ids = np.arange(n_max, dtype=np.int32)
    x = rng.uniform(0.0, NUD_NX, n_max)
    x[x == 0.0] = 0.5
    y = rng.uniform(1.0, NUD_NY, n_max)
    truth.init_particles(ids, x, y, tau_p=np.zeros(n_max))
    T_at = truth.sample("T", x, y)
    truth.collect_records()
    truth.step(steps)
    T_end = truth.fields()["T"]
    truth.step(NUD_EVERY)  # the record the last interval reads
    recs = {r["step"] - spinup: {"id": r["id"], "x": r["x"], "y": r["y"], "temp": r["temp"]}
            for r in truth.records}
    truth.close()
    l2, ms = [], []
    for n in NUD_SWEEP_OBS:
        sub = {k: {f: v[f][:n] for f in v} for k, v in recs.items()}
        row = []
        for alpha in NUD_SWEEP_ALPHA:
            s = nudged_solver(rbn, pt, sub, ids[:n], x[:n], y[:n], T_at[:n], T_first, rbn.RBN_FP32, alpha,
                              rbn.RBN_DEPOSIT_BILINEAR)
            # the first record interval untimed (it uploads the observation table), the rest
            # timed: `steps` steps in all
            s.step(NUD_EVERY)
            s.synchronize()
            t0 = time.perf_counter()
            s.step(steps - NUD_EVERY)
            s.synchronize()
            dt = time.perf_counter() - t0
            row.append(float(f"{np.sqrt(np.mean((s.fields()['T'] - T_end) ** 2)):.4g}"))
            if alpha == NUD_ALPHA:
                ms.append(round(1e3 * dt / (steps - NUD_EVERY), 4))
            s.close()
        l2.append(row)
    return {"steps": steps, "n_obs": list(NUD_SWEEP_OBS), "alpha": list(NUD_SWEEP_ALPHA),
            "l2_T_error_initial": float(f"{np.sqrt(np.mean((T_first - T_start) ** 2)):.4g}"),
            "l2_T_error_final": l2, "ms_per_step_alpha_0.1": ms}


def run_nudging(rbn, steps, spinup, chunks=10):
    """Nudging configuration: a truth run at 1024 x 512, Ra = 1e7, spun up from the
    conduction profile with four seeded convection rolls and 1e-3 noise (seed 5) into
    developed convection, then followed for
    `steps` steps with 1e4 tracers recorded every 5 steps; nudged runs start from the
    conduction profile with independent noise (seed 6) and are nudged toward the recorded
    temperatures along the same trajectories (alpha = 0.1).  Reported: the nudged FP32 step
    rate (bilinear deposit), the L2 temperature error against the truth every steps/chunks
    steps for the bilinear and the 7x7 deposit and for a free (alpha = 0) run, and the
    FP32-vs-FP64 temperature difference of the bilinear nudged run at the end."""
    pt, truth, T_first = nudging_truth(rbn, spinup)
    rng = np.random.default_rng(7)
    ids = np.arange(NUD_OBS, dtype=np.int32)
    x = rng.uniform(0.0, NUD_NX, NUD_OBS)
    x[x == 0.0] = 0.5
    y = rng.uniform(1.0, NUD_NY, NUD_OBS)
    chunk = max(steps // chunks, NUD_EVERY)
    e_spin = truth.energy()
    T_start = truth.fields()["T"]
    truth.init_particles(ids, x, y, tau_p=np.zeros(NUD_OBS))
    T_at = truth.sample("T", x, y)
    truth.collect_records()
    T_ref = []
    for _ in range(chunks):
        truth.step(chunk)
        T_ref.append(truth.fields()["T"])
    truth.step(NUD_EVERY)  # the record the last interval reads
    recs = {r["step"] - spinup: {"id": r["id"], "x": r["x"], "y": r["y"], "temp": r["temp"]}
            for r in truth.records}
    truth.close()

    def nudged(prec, alpha, deposit):
        return nudged_solver(rbn, pt, recs, ids, x, y, T_at, T_first, prec, alpha, deposit)

    def l2_curve(prec, alpha, deposit):
        s = nudged(prec, alpha, deposit)
        out, T = [], None
        for k in range(chunks):
            s.step(chunk)
            T = s.fields()["T"]
            out.append(float(f"{np.sqrt(np.mean((T - T_ref[k]) ** 2)):.4g}"))
        s.close()
        return out, T

    bil = rbn.RBN_DEPOSIT_BILINEAR
    s = nudged(rbn.RBN_FP32, NUD_ALPHA, bil)
    s.step(NUD_EVERY * 4)
    s.synchronize()
    n = chunk * chunks - NUD_EVERY * 4
    t0 = time.perf_counter()
    s.step(n)
    s.synchronize()
    dt = time.perf_counter() - t0
    s.close()
    e32, T32 = l2_curve(rbn.RBN_FP32, NUD_ALPHA, bil)
    e64, T64 = l2_curve(rbn.RBN_FP64, NUD_ALPHA, bil)
    ebox, _ = l2_curve(rbn.RBN_FP32, NUD_ALPHA, rbn.RBN_DEPOSIT_BOX7)
    efree, _ = l2_curve(rbn.RBN_FP32, 0.0, bil)
    return {"config": f"{NUD_NX}x{NUD_NY}, Ra={NUD_RA:g}, truth spun up {spinup} steps, {NUD_OBS} tracer "
                      f"observations every {NUD_EVERY} steps, alpha={NUD_ALPHA}, {chunk * chunks} steps",
            "mlups_fp32_bilinear": round(NUD_NX * NUD_NY * n / dt / 1e6, 1),
            "ms_per_step_fp32": round(1e3 * dt / n, 4),
            "truth_energy_T_at_start": [float(f"{v:.4g}") for v in e_spin],
            "l2_T_error_initial": float(f"{np.sqrt(np.mean((T_first - T_start) ** 2)):.4g}"),
            "every_steps": chunk,
            "l2_T_error_bilinear_fp32": e32,
            "l2_T_error_bilinear_fp64": e64,
            "l2_T_error_box7_fp32": ebox,
            "l2_T_error_free_fp32": efree,
            "fp32_vs_fp64_max_dT": float(f"{np.max(np.abs(T32 - T64)):.3g}")}


ADJ_NX, ADJ_NY, ADJ_RA = 512, 256, 1e7
ADJ_OBS, ADJ_EVERY, ADJ_ALPHA = 10_000, 10, 0.1


def adjoint_bytes_per_node(precision):
    # populations entering the step (18), lagged force, nud_term, pulled adjoints of the next
    # step (18), force and nud_term adjoints: loads; pulled adjoints (18) + 2: stores; plus
    # the deposit count (int) and mean (double) of the step.
    r = 8 if precision == 0 else 4
    return (18 + 2 + 18 + 2) * r + (18 + 2) * r + 4 + 8


def run_adjoint(rbn, window, ckpt):
    """Adjoint configuration: 512 x 256, Ra = 1e7, FP32.  A truth run with 1e4 tracers
    recorded every 10 steps provides the observations; the nudged run starts from another
    initial temperature.  Timed: the plain nudged forward (rbn_step) over the window, then
    the checkpointed forward (rbn_adj_forward) and the reverse sweep (rbn_adj_backward)."""
    pt = rbn.default_params(nx=ADJ_NX, ny=ADJ_NY, Tup=0.0, Tdown=1.0, scratch=1, nstart=0, nout=10**9,
                            noutconfig=ADJ_EVERY, prtcl_writefull=0, particle=1,
                            precision=rbn.RBN_FP32, T0_shift=0.5, rho0_shift=1.0)
    pt = rbn.params_from_ra(pt, ADJ_RA, PR, U_FF)
    rng = np.random.default_rng(21)
    lin = np.tile(pt.Tdown + (pt.Tup - pt.Tdown) * np.arange(ADJ_NY) / (ADJ_NY - 1), (ADJ_NX, 1))
    T_truth = lin + rng.uniform(-1e-2, 1e-2, lin.shape)
    T_first = lin + rng.uniform(-1e-2, 1e-2, lin.shape)
    ids = np.arange(ADJ_OBS, dtype=np.int32)
    x = rng.uniform(0.0, ADJ_NX, ADJ_OBS)
    x[x == 0.0] = 0.5
    y = rng.uniform(1.0, ADJ_NY, ADJ_OBS)
    truth = rbn.Solver(pt, 0)
    truth.init_fields(T=T_truth)
    truth.init_particles(ids, x, y, tau_p=np.zeros(ADJ_OBS))
    T_at = truth.sample("T", x, y)
    truth.collect_records()
    truth.step(window + ADJ_EVERY)
    recs = {r["step"]: {"id": r["id"], "x": r["x"], "y": r["y"], "temp": r["temp"]} for r in truth.records}
    truth.close()
    pn = pt.copy()
    pn.variant = rbn.RBN_VARIANT_NUDGING
    pn.ntempread = ADJ_EVERY
    pn.alpha = ADJ_ALPHA

    def fresh():
        s = rbn.Solver(pn, 0)
        s.set_observation_source(recs)
        s.init_fields(T=T_first)
        s.init_observers(ids, x, y, T_at)
        return s

    s = fresh()
    s.step(window // 10)  # warm-up
    s.close()
    s = fresh()
    t0 = time.perf_counter()
    s.step(window)
    s.synchronize()
    t_plain = time.perf_counter() - t0
    s.close()
    a = fresh()
    a.adj_forward(window, ckpt)  # warm-up (allocations)
    a.adj_backward()
    a.init_fields(T=T_first)
    a.init_observers(ids, x, y, T_at)
    t0 = time.perf_counter()
    J = a.adj_forward(window, ckpt)
    t1 = time.perf_counter()
    g, ga = a.adj_backward()
    t2 = time.perf_counter()
    a.kernel_timing(True)
    a.kernel_stats("adjoint", reset=True)
    a.adj_backward()
    n_adj, ms_adj = a.kernel_stats("adjoint")
    a.kernel_timing(False)
    a.close()
    avg = ms_adj / max(n_adj, 1)
    nodes = ADJ_NX * ADJ_NY
    return {"config": f"{ADJ_NX}x{ADJ_NY}, Ra={ADJ_RA:g}, FP32, window {window} steps, checkpoint every {ckpt}, "
                      f"{ADJ_OBS} tracer observations every {ADJ_EVERY} steps, alpha={ADJ_ALPHA}",
            "ms_per_step_plain": round(1e3 * t_plain / window, 4),
            "ms_per_step_forward_ckpt": round(1e3 * (t1 - t0) / window, 4),
            "ms_per_step_backward": round(1e3 * (t2 - t1) / window, 4),
            "cost_vs_plain_forward": round((t2 - t0) / t_plain, 3),
            "adjoint_kernel_avg_ms": round(avg, 5),
            "adjoint_kernel_gbps": round(nodes * adjoint_bytes_per_node(1) / (avg * 1e-3) / 1e9, 1),
            "J": J, "max_abs_dJ_dT0": float(np.max(np.abs(g))), "dJ_dalpha": ga,
            "finite": bool(np.isfinite(g).all() and np.isfinite(ga))}


def launch_bytes(precision, record_fraction, particles, steps=1):
    """Algorithmic bytes of one lattice launch of `steps` steps, averaged over the steps: the
    lattice bytes of every step, T on record steps, and with particles=True as many carried
    particle updates.  A two-step launch keeps the first step's populations in LDS, so it
    moves fewer bytes through HBM than these (its measured traffic is reported beside)."""
    s = 8 if precision == 0 else 4
    b = NX * NY * (lbm_bytes_per_node(precision) + record_fraction * s)
    return steps * (b + (particle_bytes_per_step(precision) if particles else 0))


TRAFFIC_PROFILES = {2: ROOT / "profiles" / "k_lbm2_32_4096x2048_pmc.txt",
                    3: ROOT / "profiles" / "k_lbm3w_32_4096x2048_pmc.txt"}


def measured_traffic(steps_per_launch):
    """HBM bytes per lattice launch of that many steps carrying particle updates, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (counters scaled by a copy kernel of
    known size), or None."""
    try:
        for line in TRAFFIC_PROFILES[steps_per_launch].read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_threads():
    """Host cores this process may use, at most 16 (the host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def run_cpu_baseline(rbn, steps, threads):
    """The CPU FP64 restatement of the reference loop (array-of-structs, per-direction sweeps,
    halo copies) on `threads` host cores, decomposed like the reference's MPI run: x-slabs of
    nx/threads columns, one per thread, each an independent restatement instance stepping its
    slab and the particles inside it (the oracle releases the GIL inside each step).  The
    slabs wrap periodically onto themselves instead of exchanging halos, so this leaves out
    the reference's halo traffic: an upper bound on its throughput on this host."""
    import threading

    from oracle.oracle import Oracle

    p = make_params(rbn, 0)
    T0, (ids, x, y, tau, zeros) = make_inputs(p, NX, NY, N_TRACER, N_INERTIAL)
    w = NX // threads
    sims = []
    placed = 0
    for k in range(threads):
        q = p.copy()
        q.nx = w
        o = Oracle(q)
        o.set_faithful(True)
        o.init_fields(T=np.ascontiguousarray(T0[k * w:(k + 1) * w]))
        m = (x >= k * w) & (x < (k + 1) * w)
        xl = x[m] - k * w
        xl[xl == 0.0] = 0.5
        o.init_particles(ids[m], xl, y[m], tau_p=tau[m], up=zeros[m], vp=zeros[m])
        placed += int(m.sum())
        sims.append(o)
    start = threading.Barrier(threads + 1)
    done = [0.0] * threads

    def work(k):
        sims[k].step(1)
        start.wait()
        sims[k].step(steps)
        done[k] = time.perf_counter()

    ts = [threading.Thread(target=work, args=(k,)) for k in range(threads)]
    for t in ts:
        t.start()
    start.wait()
    t0 = time.perf_counter()
    for t in ts:
        t.join()
    dt = max(done) - t0
    for o in sims:
        o.close()
    return {"value": w * threads * NY * steps / dt / 1e6, "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"oracle/rbn_oracle.c (CPU FP64 restatement of the reference truth-run loop, "
                      f"cost-faithful mode) on {threads} threads, one x-slab of {w}x{NY} each with the "
                      f"particles inside it ({NX}x{NY} + {N_TRACER}+{N_INERTIAL} particles in total), "
                      f"{steps} steps after 1 warm-up step, slabs periodic on themselves (no halo exchange); "
                      f"host {platform.processor() or platform.machine()}, nproc {os.cpu_count()}",
            "seconds": dt, "particles": placed}


def finite_or_null(v):
    """The result with every non-finite float replaced by null (a diverged run), so that the
    line stays strict JSON."""
    if isinstance(v, dict):
        return {k: finite_or_null(x) for k, x in v.items()}
    if isinstance(v, (list, tuple)):
        return [finite_or_null(x) for x in v]
    if isinstance(v, float) and not np.isfinite(v):
        return None
    return v


# The headline is timed over at least this many steps after at least MIN_WARMUP warm-up
# steps, whatever --steps / --warmup ask for (a steady window of the benchmark workload).
MIN_TIMED_STEPS, MIN_WARMUP = 1000, 50
# Ensemble legs: at least this many timed ensemble steps (ten record cycles of 30 steps).
MIN_ENS_STEPS = 300


def launch_min_bytes(precision, record_fraction, particles=True, steps=2):
    """The fewest HBM bytes a lattice launch of `steps` steps can move: every population and
    the force read once (19 reals per node) and written once (19), the velocity of every step
    written (2 reals each), rho and T written on the launches holding a record step (2 reals,
    on steps * record_fraction of the launches), plus the carried particle
    updates' algorithmic bytes (one per step)."""
    s = 8 if precision == 0 else 4
    b = NX * NY * s * (38 + 2 * steps + 2 * steps * record_fraction)
    return b + (steps * particle_bytes_per_step(precision) if particles else 0)


def run_leg(name, fn, *args, **kw):
    """Run one benchmark leg; a failure is reported as {"error": ...} inside the JSON line
    (with the traceback on stderr) instead of ending the run."""
    import traceback

    t0 = time.perf_counter()
    try:
        r = fn(*args, **kw)
    except Exception as exc:  # noqa: BLE001 - every leg failure is reported, none is fatal
        traceback.print_exc(file=sys.stderr)
        print(f"bench: {name} failed after {time.perf_counter() - t0:.1f} s: {exc!r}", file=sys.stderr, flush=True)
        return {"error": f"{type(exc).__name__}: {exc}"}
    print(f"bench: {name} done in {time.perf_counter() - t0:.1f} s {r}", file=sys.stderr, flush=True)
    return r


def ok(r):
    return isinstance(r, dict) and "error" not in r


def lattice_summary(r, kernel_key, bytes_launch, extra=None):
    """MLUPS, ms per step and the dominant launch's roofline fraction of one lattice leg."""
    if not ok(r):
        return r
    out = {"mlups": round(r["mlups"], 1), "ms_per_step": round(r["ms_per_step"], 5), "steps": r["steps"]}
    if kernel_key:
        avg = r[kernel_key]
        out["avg_launch_ms"] = round(avg, 5)
        out["bytes_per_launch"] = round(bytes_launch)
        out["roofline_frac"] = round(bytes_launch / (avg * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4) if avg > 0 else None
    out.update(extra or {})
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000,
                    help=f"timed steps of the headline (at least {MIN_TIMED_STEPS} are timed)")
    ap.add_argument("--warmup", type=int, default=100, help=f"warm-up steps (at least {MIN_WARMUP})")
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU baseline threads (default: host cores, at most 16)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 companion run")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble run")
    ap.add_argument("--ensemble-tuning", type=str, default="",
                    help="tile_w,tile_h,wg_per_cu,order for the nudged ensemble instead of autotuning")
    ap.add_argument("--ensemble-kind", choices=("both", "nudged", "truth"), default="both",
                    help="ensemble legs to run: nudged members (the ensemble configuration), truth-run members, or both")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-window run")
    ap.add_argument("--no-nudging", action="store_true", help="skip the nudging run")
    ap.add_argument("--nud-steps", type=int, default=20000)
    ap.add_argument("--nud-sweep-steps", type=int, default=2000, help="steps of each N_p x alpha sweep run (0: skip)")
    ap.add_argument("--nud-spinup", type=int, default=30000)
    ap.add_argument("--adj-window", type=int, default=2000)
    ap.add_argument("--adj-ckpt", type=int, default=100)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the solver runs on one GPU; an x-slab decomposition over several GPUs "
              "is not implemented. Run with --gpus 1.", file=sys.stderr)
        return 2
    from lagrangianrbnudging_amd import rbn

    if rbn.device_count() == 0:
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1
    steps = max(args.steps, MIN_TIMED_STEPS)
    warmup = max(args.warmup, MIN_WARMUP)
    half, half_w = max(steps // 2, MIN_TIMED_STEPS // 2), max(warmup // 2, MIN_WARMUP)
    print(f"bench: fp32 {NX}x{NY} steps={steps} warmup={warmup}", file=sys.stderr, flush=True)
    r32 = run_leg("fp32", run_gpu, rbn, rbn.RBN_FP32, steps, warmup)
    r64 = None if args.no_fp64 else run_leg("fp64", run_gpu, rbn, rbn.RBN_FP64, half, half_w)
    two = run_leg("fp32 two steps per launch", run_gpu, rbn, rbn.RBN_FP32, half, half_w, lbm_steps=2)
    one = run_leg("fp32 one step per launch", run_gpu, rbn, rbn.RBN_FP32, half, half_w, lbm_steps=1)
    sep = run_leg("fp32 separate particle launches", run_gpu, rbn, rbn.RBN_FP32, half, half_w, particle_fuse=0)
    host = run_leg("fp32 records to host", run_gpu, rbn, rbn.RBN_FP32, steps, warmup, timing=False,
                   records_to_host=True)
    ens = None
    if not args.no_ensemble:
        tun = tuple(int(v) for v in args.ensemble_tuning.split(",")) if args.ensemble_tuning else None
        ens_steps = max(args.steps // 5, MIN_ENS_STEPS)
        ens = {}
        if args.ensemble_kind != "truth":
            ens = run_leg("ensemble (nudged members)", run_ensemble, rbn, ens_steps, 20, tun, nudged=True)
        if args.ensemble_kind != "nudged":
            ens["truth_run_members"] = run_leg("ensemble (truth-run members)", run_ensemble, rbn, ens_steps, 20,
                                               None, nudged=False)
    nud = None
    if not args.no_nudging:
        nud = run_leg("nudging", run_nudging, rbn, args.nud_steps, args.nud_spinup)
        if args.nud_sweep_steps > 0:
            nud["sweep"] = run_leg("nudging sweep", run_nudging_sweep, rbn, args.nud_sweep_steps, args.nud_spinup)
    adj = None if args.no_adjoint else run_leg("adjoint", run_adjoint, rbn, args.adj_window, args.adj_ckpt)
    cpu = None
    if not args.no_cpu:
        cpu = run_leg("cpu", run_cpu_baseline, rbn, args.cpu_steps, args.cpu_threads or cpu_threads())

    line = {
        "metric": METRIC,
        "value": round(r32["mlups"], 1) if ok(r32) else None,
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": r32["steps"] if ok(r32) else steps,
        "warmup": warmup,
        "steps_requested": args.steps,
        "ms_per_step": round(r32["ms_per_step"], 5) if ok(r32) else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"plain RB step {NX}x{NY}, Ra=1e8, Pr=1, FP32 shifted populations, "
                        f"{N_TRACER} tracers + {N_INERTIAL} inertial (St in [0.01,1]), records every {RECORD_EVERY} steps",
            "nx": NX, "ny": NY, "Ra": RA, "Pr": PR, "particles": N_TRACER + N_INERTIAL,
            "record_every": RECORD_EVERY,
        },
        "mlups_fp64": round(r64["mlups"], 1) if ok(r64) else None,
        "ms_per_step_fp64": round(r64["ms_per_step"], 5) if ok(r64) else None,
    }
    if ok(r32):
        # the dominant launch: three steps (where record steps allow) or two; record steps end
        # launches, so three-step launches never write the moments at 20-step records
        spl = 3 if r32["lbm3_timed"] > 0 else 2
        rf = r32["record_fraction"]
        bytes_launch = launch_bytes(1, rf, True, steps=spl)
        min_bytes = launch_min_bytes(1, rf, steps=spl)
        avg = r32[f"lbm{spl}_avg_ms"]
        achieved = bytes_launch / (avg * 1e-3) / 1e9
        traffic = measured_traffic(spl)
        line["roofline"] = {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": traffic,
            "traffic_frac": traffic and round(traffic / (avg * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
            "min_traffic_bytes": round(min_bytes),
            "min_traffic_frac": round(min_bytes / (avg * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
            "kernel": ("k_lbm3w<float, 256> (three fused f/g collide+stream steps by row sweeps, the first two "
                       "steps' populations passed through LDS, + the previous launch's particle updates)") if spl == 3
            else ("k_lbm2<float, 8> (two fused f/g collide+stream steps, the first step's populations in "
                  "LDS, + the two previous steps' particle updates)"),
            "steps_per_launch": spl,
            "bytes_per_launch": round(bytes_launch),
            "avg_launch_ms": round(avg, 5),
            "launches_timed": r32[f"lbm{spl}_timed"],
            "timed_every": TIMING_EVERY,
            "note": "achieved counts every step's algorithmic bytes (160 B/LUP each), so frac can exceed 1 when "
                    "the earlier steps of a launch stay on chip; min_traffic_frac divides the fewest bytes such a "
                    "launch can move (38 + 2 x steps reals per node + particles) by the same duration and "
                    "cannot exceed 1",
        }
    else:
        line["roofline"] = None
        line["error"] = r32.get("error")
    line["cpu_baseline"] = cpu and ({k: (round(v, 3) if isinstance(v, float) else v) for k, v in cpu.items()}
                                    if ok(cpu) else cpu)
    line["fp64"] = None if r64 is None else lattice_summary(
        r64, "lbm2_avg_ms", launch_bytes(0, r64.get("record_fraction", 0.0), True, steps=2),
        {"min_traffic_frac": ok(r64) and r64["lbm2_avg_ms"] > 0 and round(
            launch_min_bytes(0, r64["record_fraction"]) / (r64["lbm2_avg_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)})
    line["records_to_host"] = lattice_summary(host, None, 0, {
        "note": "PCIe-inclusive: every record also copied to pinned host memory and handed to a callback"})
    if ok(two):
        line["two_steps_per_launch"] = lattice_summary(
            two, "lbm2_avg_ms", launch_bytes(1, two["record_fraction"], True, steps=2),
            {"lattice_kernel": "k_lbm2<float, 8> (64 x 8 tiles, two steps + the previous two steps' particle updates)",
             "min_traffic_frac": two["lbm2_avg_ms"] > 0 and round(
                 launch_min_bytes(1, two["record_fraction"]) / (two["lbm2_avg_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)})
    else:
        line["two_steps_per_launch"] = two
    line["one_step_per_launch"] = lattice_summary(
        one, "fused_avg_ms", launch_bytes(1, one.get("record_fraction", 0.0), True),
        {"lattice_kernel": "k_lbm_fused32 (one step + the previous step's particle update)"})
    if ok(sep):
        sep_bytes = launch_bytes(1, sep["record_fraction"], False)
        line["separate_particle_launches"] = lattice_summary(
            sep, "lbm_avg_ms", sep_bytes, {"lattice_kernel": "k_lbm32", "particles_avg_ms": round(sep["particles_avg_ms"], 5)})
    else:
        line["separate_particle_launches"] = sep
    line["ensemble"] = ens
    line["nudging"] = nud
    line["adjoint"] = adj
    line["finite"] = all(r["finite"] for r in (r32, r64, adj) if ok(r))
    print(json.dumps(finite_or_null(line), allow_nan=False), flush=True)
    return 0 if ok(r32) else 1


if __name__ == "__main__":
    sys.exit(main())
