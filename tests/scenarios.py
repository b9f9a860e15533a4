"""Seeded scenario builders shared by the tests: the same inputs go to librbn and the oracle."""
from __future__ import annotations

import numpy as np

from lagrangianrbnudging_amd import rbn


def rb_params(nx, ny, Ra, Pr=1.0, u_ff=0.05, **kw) -> rbn.Params:
    """Rayleigh–Bénard parameters with Tdown = 1 (bottom), Tup = 0 (top)."""
    base = dict(nx=nx, ny=ny, Tup=0.0, Tdown=1.0, scratch=1, nstart=0, nout=100000,
                noutconfig=10, prtcl_writefull=0, particle=1, T0_shift=0.5)
    base.update(kw)
    p = rbn.default_params(**{k: v for k, v in base.items() if k not in ("relax", "relaxg", "gravity")})
    p = rbn.params_from_ra(p, Ra, Pr, u_ff)
    for k in ("relax", "relaxg", "gravity"):
        if k in kw:
            setattr(p, k, kw[k])
    return p


def conduction_T(p, amp=1e-3, seed=1) -> np.ndarray:
    """Linear conduction profile plus seeded uniform noise, [nx][ny]."""
    rng = np.random.default_rng(seed)
    j = np.arange(p.ny)
    lin = p.Tdown + (p.Tup - p.Tdown) * j / (p.ny - 1)
    return np.tile(lin, (p.nx, 1)) + rng.uniform(-amp, amp, (p.nx, p.ny))


def random_particles(p, n_tracer, n_inertial, seed=2, tau_range=(2.0, 200.0)):
    """Uniform random positions in (0, nx) x [1, ny]; inertial tau log-uniform in tau_range."""
    rng = np.random.default_rng(seed)
    n = n_tracer + n_inertial
    x = rng.uniform(0.0, p.nx, n)
    x[x == 0.0] = 0.5
    y = rng.uniform(1.0, p.ny, n)
    tau = np.zeros(n)
    lo, hi = np.log10(tau_range[0]), np.log10(tau_range[1])
    tau[n_tracer:] = 10.0 ** rng.uniform(lo, hi, n_inertial)
    ids = rng.permutation(n).astype(np.int32)
    return ids, x, y, tau


def setup(sim, p, T0, particles=None):
    sim.init_fields(T=T0)
    if particles is not None:
        ids, x, y, tau = particles
        sim.init_particles(ids, x, y, tau_p=tau)
    return sim


def truth_records(p, T0, particles, steps, device=0):
    """Run a truth run on the GPU and return its records keyed by step."""
    s = rbn.Solver(p, device)
    s.collect_records()
    setup(s, p, T0, particles)
    s.step(steps)
    recs = {r["step"]: r for r in s.records}
    s.close()
    return recs


def synthetic_observations(nx, ny, n, steps, every, seed):
    """Observers on straight lines that cross the periodic x boundary and the walls."""
    rng = np.random.default_rng(seed)
    x0 = np.concatenate([[0.3, 3.2, nx - 2.0, nx - 1.4, nx - 0.2, 1.9], rng.uniform(0, nx, n - 6)])
    y0 = np.concatenate([[1.1, ny - 0.1, 2.4, 1.6, ny - 1.2, ny / 2], rng.uniform(1, ny, n - 6)])
    vx = rng.uniform(-0.15, 0.15, n)
    vy = rng.uniform(-0.1, 0.1, n)
    T = rng.uniform(0, 1, n)
    ids = np.arange(n, dtype=np.int32)
    recs = {}
    for s in range(every, steps + every, every):
        xs = x0 + vx * s
        ys = y0 + vy * s
        xs = np.mod(xs, nx)
        ys = 1 + np.mod(ys - 1, ny - 1)
        perm = rng.permutation(n)
        recs[s] = {"id": ids[perm], "x": xs[perm], "y": ys[perm], "temp": (T + 0.01 * s)[perm]}
    return ids, x0, y0, T, recs


def assert_equal_arrays(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    if not np.array_equal(a, b):
        diff = np.abs(a.astype(np.float64) - b.astype(np.float64))
        k = np.unravel_index(np.argmax(diff), diff.shape)
        raise AssertionError(f"{what}: {np.count_nonzero(a != b)} of {a.size} differ; max |diff| {diff[k]:.3e} at {k} "
                             f"({a[k]!r} vs {b[k]!r})")


def max_rel(a, b, floor=1e-300):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), floor))
