"""Complex-step tangent-linear reference for the misfit of the nudged run.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the FP64 tangent-linear reference of the
adjoint window (rbn_adj_forward / rbn_adj_backward).  The reference has no gradient, so
this is a restatement of the nudged step (Nudging/code/tempvel_nudge.c: pull streaming with
the f half-way bounce-back :757-894 and the zero-gradient g walls :684-702, hydrovar
:1069-1102, the lagged buoyancy force :744-755, equilif/equilig :1097-1180, collisf
:1314-1328 and the nudged collisg :1917-1945) written with numpy on complex128 fields.

Evaluating the misfit at (T0 + i h dT, alpha + i h da) with h = 1e-30 gives
    Im J / h = dJ/dT0 . dT + dJ/dalpha * da
to working precision (complex-step differentiation: no difference of nearly equal numbers),
which is the directional derivative a tangent-linear model computes.

What the step does not differentiate is data: the deposit counts and running means of each
step and the observer positions / temperatures at the observation steps depend only on the
observation records.  They are taken from the C oracle (oracle/rbn_oracle.c) run step by
step on the same inputs, and the restatement's real part is pinned against that oracle's
populations in tests/test_tlm_reference.py.
"""
from __future__ import annotations

import numpy as np

from .oracle import Oracle

W = np.array([16.0 / 36] + [4.0 / 36] * 4 + [1.0 / 36] * 4)
CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
OPP = [0, 3, 4, 1, 2, 7, 8, 5, 6]


def nudging_data(p, T0, obs, records, nsteps):
    """Per step of a nudged oracle run: deposit counts/means (and weights of the bilinear
    deposit) and, at observation steps
    (istep % ntempread == 0), the observer positions and temperatures (observer order) and
    the oracle's temperature field; also the oracle's final populations."""
    o = Oracle(p)
    try:
        o.init_fields(T=T0)
        o.init_observers(*obs)
        o.set_observation_source(records)
        out = []
        for _ in range(nsteps):
            o.step(1)
            nd = o.nudging()
            d = {"nvol": nd["nudge_vol"].copy(), "tbar": nd["tempreadint"].copy(),
                 "w": nd["weight"].copy() if p.deposit == 1 else None}
            if o.istep % p.ntempread == 0:
                s = o.observer_state()
                d["obs"] = (s["xint"], s["yint"], s["Tint"])
                d["T"] = o.fields()["T"]  # the oracle's own temperature of the step
            out.append(d)
        f, g = o.populations()
        return out, (f, g)
    finally:
        o.close()


def bilinear(T, x, y):
    """Bilinear interpolation at halo coordinates (1-based nodes, periodic ghost copies in x
    and y, C truncation of the corner index), as the trajectory records use it."""
    nx, ny = T.shape
    x1 = np.trunc(x).astype(np.int64)
    y1 = np.trunc(y).astype(np.int64)
    x2, y2 = x1 + 1, y1 + 1

    def at(ic, jc):
        I = np.where(ic == 0, nx, np.where(ic == nx + 1, 1, ic))
        J = np.where(jc == 0, ny, np.where(jc == ny + 1, 1, jc))
        return T[I - 1, J - 1]

    return (at(x1, y1) * (x2 - x) * (y2 - y) + at(x2, y1) * (x - x1) * (y2 - y)
            + at(x1, y2) * (x2 - x) * (y - y1) + at(x2, y2) * (x - x1) * (y - y1))


class NudgedStep:
    """The nudged lattice step on (9, nx, ny) population arrays of any numpy dtype."""

    def __init__(self, p):
        self.nx, self.ny = p.nx, p.ny
        self.om = 1.0 / p.relax
        self.omg = 1.0 / p.relaxg
        self.g0 = -p.gravity
        self.src = (1.0 - 0.5 * self.om) * W
        self.bw = (1.0 - 0.5 * self.omg) * W
        self.rho0 = p.rhom1

    def init(self, T0):
        """Equilibrium at rest with density rhom1 and temperature T0 (rbn_init_fields)."""
        T0 = np.asarray(T0)
        rho = np.full(T0.shape, self.rho0, dtype=T0.dtype)
        f = W[:, None, None] * rho[None]
        g = W[:, None, None] * T0[None]
        return f, g, self.g0 * rho * T0, np.zeros_like(T0)

    def pull_f(self, f):
        out = np.empty_like(f)
        for k in range(9):
            sh = np.roll(f[k], CX[k], axis=0)  # out[i] = f[i - cx]
            if CY[k] == 0:
                out[k] = sh
            elif CY[k] == 1:
                out[k][:, 1:] = sh[:, :-1]
                out[k][:, 0] = f[OPP[k]][:, 0]  # bounce-back at the bottom wall
            else:
                out[k][:, :-1] = sh[:, 1:]
                out[k][:, -1] = f[OPP[k]][:, -1]  # bounce-back at the top wall
        return out

    def pull_g(self, g):
        ny = self.ny
        out = np.empty_like(g)
        for k in range(9):
            sh = np.roll(g[k], CX[k], axis=0)
            rows = np.clip(np.arange(ny) - CY[k], 1, ny - 2)  # wall rows replaced by their copies
            out[k] = sh[:, rows]
            out[k][:, 0] = g[k][:, 1]
            out[k][:, ny - 1] = g[k][:, ny - 2]
        return out

    def step(self, state, alpha, data):
        f, g, F, nud = state
        fp, gp = self.pull_f(f), self.pull_g(g)
        nv = data["nvol"] > 0
        rho, S = fp.sum(0), gp.sum(0)
        T = np.where(nv, S + 0.5 * nud, S)
        jx = np.tensordot(CX, fp, axes=1)
        jy = np.tensordot(CY, fp, axes=1)
        u = jx / rho
        v = (jy + 0.5 * F) / rho
        usq = u * u + v * v
        Fn = self.g0 * rho * T
        w = data.get("w")
        Q = alpha * (T - data["tbar"]) if w is None else alpha * (np.minimum(w, 1.0) * (T - data["tbar"]))
        fo, go = np.empty_like(fp), np.empty_like(gp)
        for k in range(9):
            cu = CX[k] * u + CY[k] * v
            b = 1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq
            fac = 3.0 * (CY[k] - v) * F + 9.0 * cu * CY[k] * F
            fo[k] = (1.0 - self.om) * fp[k] + self.om * W[k] * rho * b + self.src[k] * fac
            go[k] = (1.0 - self.omg) * gp[k] + self.omg * W[k] * T * b - np.where(nv, self.bw[k] * Q, 0.0)
        return (fo, go, Fn, np.where(nv, -Q, nud)), T


def misfit(p, T0, alpha, data):
    """J over the window described by data (one entry per step); also the final state."""
    model = NudgedStep(p)
    state = model.init(T0)
    J = 0.0
    for d in data:
        state, T = model.step(state, alpha, d)
        if "obs" in d:
            x, y, To = d["obs"]
            r = bilinear(T, x, y) - To
            J = J + (r * r).sum()
    return J, state


def directional_derivative(p, T0, alpha, data, dT, dalpha=0.0, h=1e-30):
    """(J, dJ/dT0 . dT + dJ/dalpha * dalpha) by complex-step differentiation."""
    Jc, _ = misfit(p, np.asarray(T0, dtype=np.complex128) + 1j * h * np.asarray(dT),
                   complex(alpha, h * dalpha), data)
    return float(Jc.real), float(Jc.imag / h)
