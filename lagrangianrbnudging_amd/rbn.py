"""ctypes binding of librbn's C-ABI (include/rbn.h).

This is the binding a Python caller of the drop-in library uses (the tests and the benchmark
use it); it adds nothing to the computation.  There is no CPU fallback: loading fails loudly
when librbn.so has not been built, and creating a solver fails when no GPU is visible.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Callable, Optional

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "librbn.so"

RBN_VARIANT_TRAJECTORIES = 0
RBN_VARIANT_NUDGING = 1
RBN_FP64 = 0
RBN_FP32 = 1
RBN_DEPOSIT_BOX7 = 0
RBN_DEPOSIT_BILINEAR = 1

STATUS = {
    0: "RBN_OK", -1: "RBN_EINVAL", -2: "RBN_EIO", -3: "RBN_EPARSE", -4: "RBN_EDEVICE",
    -5: "RBN_ESTATE", -6: "RBN_ERANGE", -7: "RBN_ENOMEM", -8: "RBN_ENUMERIC",
    -9: "RBN_ENOTFOUND", -10: "RBN_ECALLBACK",
}
STATUS_CODE = {name: code for code, name in STATUS.items()}


class RbnError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    """Mirror of rbn_params (field order and types as in include/rbn.h)."""

    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("scratch", C.c_int), ("contin", C.c_int),
        ("nstart", C.c_int), ("nsteps", C.c_int), ("nout", C.c_int), ("noutconfig", C.c_int),
        ("ntempread", C.c_int),
        ("rhom1", C.c_double), ("relax", C.c_double), ("relaxg", C.c_double), ("Tup", C.c_double),
        ("Tdown", C.c_double), ("gravity", C.c_double), ("alpha", C.c_double),
        ("particle", C.c_int), ("particle_scratch", C.c_int), ("nprtcl", C.c_int),
        ("tracer", C.c_int), ("heavyprtcl", C.c_int),
        ("taulag", C.c_double), ("beta", C.c_double),
        ("prtcl_writefull", C.c_int),
        ("variant", C.c_int), ("precision", C.c_int), ("deposit", C.c_int), ("fix_q3", C.c_int),
        ("ndiag", C.c_int),
        ("T0_shift", C.c_double), ("rho0_shift", C.c_double), ("heavy_init_factor", C.c_double),
    ]

    def copy(self) -> "Params":
        p = Params()
        C.memmove(C.byref(p), C.byref(self), C.sizeof(Params))
        return p

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Particles(C.Structure):
    _fields_ = [("n", C.c_int), ("id", C.POINTER(C.c_int))] + [
        (k, C.POINTER(C.c_double)) for k in ("x", "y", "u", "v", "xo", "yo", "uo", "vo")
    ]


_DP = C.POINTER(C.c_double)


class Record(C.Structure):
    _fields_ = [("step", C.c_int), ("n", C.c_int), ("id", C.POINTER(C.c_int))] + [
        (k, _DP) for k in ("x", "y", "u", "v", "xo", "yo", "uo", "vo", "temp",
                           "dtu", "dtv", "dxu", "dxv", "dyu", "dyv", "fu", "fv")
    ]


RECORD_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(Record))
OBS_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int), _DP, _DP, _DP)
DIAG_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_double, C.c_double)

_lib: Optional[C.CDLL] = None

# name -> (restype, argtypes)
_SIGS = {
    "rbn_version": (C.c_char_p, []),
    "rbn_api_version": (C.c_int, []),
    "rbn_last_error": (C.c_char_p, [C.c_void_p]),
    "rbn_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "rbn_params_default": (None, [C.POINTER(Params)]),
    "rbn_params_parse": (C.c_int, [C.c_char_p, C.POINTER(Params)]),
    "rbn_params_parse_text": (C.c_int, [C.c_char_p, C.POINTER(Params)]),
    "rbn_params_set": (C.c_int, [C.POINTER(Params), C.c_char_p, C.c_double]),
    "rbn_params_rayleigh": (C.c_double, [C.POINTER(Params)]),
    "rbn_params_from_ra": (C.c_int, [C.POINTER(Params), C.c_double, C.c_double, C.c_double]),
    "rbn_params_check": (C.c_int, [C.POINTER(Params)]),
    "rbn_create": (C.c_int, [C.POINTER(Params), C.c_int, C.POINTER(C.c_void_p)]),
    "rbn_destroy": (None, [C.c_void_p]),
    "rbn_get_params": (C.c_int, [C.c_void_p, C.POINTER(Params)]),
    "rbn_init_fields": (C.c_int, [C.c_void_p, _DP, _DP, _DP, _DP]),
    "rbn_init_conduction": (C.c_int, [C.c_void_p, C.c_double, C.c_ulonglong]),
    "rbn_set_populations": (C.c_int, [C.c_void_p, _DP, _DP, _DP, _DP, _DP, _DP]),
    "rbn_get_populations": (C.c_int, [C.c_void_p, _DP, _DP]),
    "rbn_get_fields": (C.c_int, [C.c_void_p, _DP, _DP, _DP, _DP]),
    "rbn_energy": (C.c_int, [C.c_void_p, _DP, _DP]),
    "rbn_profiles": (C.c_int, [C.c_void_p, _DP]),
    "rbn_sample": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _DP, _DP, _DP]),
    "rbn_sample_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _DP, _DP, _DP]),
    "rbn_init_particles": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), _DP, _DP, _DP, _DP, _DP, _DP]),
    "rbn_get_particles": (C.c_int, [C.c_void_p, C.POINTER(Particles)]),
    "rbn_num_particles": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "rbn_set_record_callback": (C.c_int, [C.c_void_p, RECORD_FN, C.c_void_p]),
    "rbn_init_observers": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), _DP, _DP, _DP]),
    "rbn_set_obs_provider": (C.c_int, [C.c_void_p, OBS_FN, C.c_void_p]),
    "rbn_set_obs_table": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "rbn_get_nudging": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), _DP, _DP]),
    "rbn_get_nudging_weights": (C.c_int, [C.c_void_p, _DP]),
    "rbn_set_diag_callback": (C.c_int, [C.c_void_p, DIAG_FN, C.c_void_p]),
    "rbn_step": (C.c_int, [C.c_void_p, C.c_int]),
    "rbn_get_step": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "rbn_synchronize": (C.c_int, [C.c_void_p]),
    "rbn_ensemble_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]),
    "rbn_ensemble_destroy": (None, [C.c_void_p]),
    "rbn_ensemble_last_error": (C.c_char_p, [C.c_void_p]),
    "rbn_ensemble_step": (C.c_int, [C.c_void_p, C.c_int]),
    "rbn_ensemble_set_tuning": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "rbn_ensemble_get_tuning": (C.c_int, [C.c_void_p] + [C.POINTER(C.c_int)] * 4),
    "rbn_ensemble_autotune": (C.c_int, [C.c_void_p, C.c_int, _DP]),
    "rbn_adj_forward": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _DP]),
    "rbn_adj_backward": (C.c_int, [C.c_void_p, _DP, _DP]),
    "rbn_set_tuning": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "rbn_get_tuning": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_int)]),
    "rbn_set_kernel_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "rbn_kernel_stats": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.POINTER(C.c_longlong), _DP]),
    "rbn_time_steps": (C.c_int, [C.c_void_p, C.c_int, _DP]),
}


def load() -> C.CDLL:
    """Load librbn.so (built in-tree by build.py); raises when it is missing."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built; run `python -m lagrangianrbnudging_amd.build`")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def declared_symbols() -> list[str]:
    return sorted(_SIGS)


def _check(rc: int, ctx=None) -> None:
    if rc != 0:
        msg = load().rbn_last_error(ctx)
        raise RbnError(rc, msg.decode() if msg else "")


def _dp(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(_DP)


def _f64(a, shape=None) -> Optional[np.ndarray]:
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != tuple(shape):
        a = a.reshape(shape)
    return a


def default_params(**overrides) -> Params:
    p = Params()
    load().rbn_params_default(C.byref(p))
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def parse_params(path: str, base: Optional[Params] = None) -> Params:
    p = base.copy() if base is not None else default_params()
    _check(load().rbn_params_parse(str(path).encode(), C.byref(p)))
    return p


def parse_params_text(text: str, base: Optional[Params] = None) -> Params:
    p = base.copy() if base is not None else default_params()
    _check(load().rbn_params_parse_text(text.encode(), C.byref(p)))
    return p


def params_from_ra(p: Params, Ra: float, Pr: float = 1.0, u_ff: float = 0.05) -> Params:
    q = p.copy()
    _check(load().rbn_params_from_ra(C.byref(q), Ra, Pr, u_ff))
    return q


def rayleigh(p: Params) -> float:
    return load().rbn_params_rayleigh(C.byref(p))


def device_count() -> int:
    n = C.c_int(0)
    _check(load().rbn_device_count(C.byref(n)))
    return n.value


RECORD_FIELDS = ("x", "y", "u", "v", "xo", "yo", "uo", "vo", "temp")
FULL_FIELDS = ("dtu", "dtv", "dxu", "dxv", "dyu", "dyv", "fu", "fv")


class Solver:
    """One lattice on one GPU (an rbn_ctx)."""

    def __init__(self, params: Params, device: int = 0):
        self.lib = load()
        self.params = params.copy()
        ctx = C.c_void_p()
        _check(self.lib.rbn_create(C.byref(self.params), device, C.byref(ctx)))
        self.ctx = ctx
        self.nx, self.ny = self.params.nx, self.params.ny
        self.records: list[dict] = []
        self.diagnostics: list[tuple[int, float, float]] = []
        self._keep = []
        self._rec_cb = None
        self._obs_cb = None
        self._diag_cb = None

    def close(self) -> None:
        if self.ctx:
            self.lib.rbn_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc: int) -> None:
        _check(rc, self.ctx)

    # -- initial state
    def init_fields(self, rho=None, T=None, u=None, v=None) -> None:
        sh = (self.nx, self.ny)
        arrs = [_f64(a, sh) for a in (rho, T, u, v)]
        self._c(self.lib.rbn_init_fields(self.ctx, *[_dp(a) for a in arrs]))

    def init_conduction(self, noise: float, seed: int) -> None:
        self._c(self.lib.rbn_init_conduction(self.ctx, noise, seed))

    def set_populations(self, f, g, rho_prev, T_prev, u_prev=None, v_prev=None) -> None:
        sh = (self.nx, self.ny)
        f = _f64(f, (self.nx, self.ny, 9))
        g = _f64(g, (self.nx, self.ny, 9))
        arrs = [_f64(a, sh) for a in (rho_prev, T_prev, u_prev, v_prev)]
        self._c(self.lib.rbn_set_populations(self.ctx, _dp(f), _dp(g), *[_dp(a) for a in arrs]))

    def populations(self) -> tuple[np.ndarray, np.ndarray]:
        f = np.empty((self.nx, self.ny, 9))
        g = np.empty((self.nx, self.ny, 9))
        self._c(self.lib.rbn_get_populations(self.ctx, _dp(f), _dp(g)))
        return f, g

    def fields(self) -> dict:
        out = {k: np.empty((self.nx, self.ny)) for k in ("rho", "T", "u", "v")}
        self._c(self.lib.rbn_get_fields(self.ctx, *[_dp(out[k]) for k in ("rho", "T", "u", "v")]))
        return out

    def energy(self) -> tuple[float, float]:
        e, t = C.c_double(), C.c_double()
        self._c(self.lib.rbn_energy(self.ctx, C.byref(e), C.byref(t)))
        return e.value, t.value

    def profiles(self) -> np.ndarray:
        """[ny][9]: x averages of rho, u, v, u^2, v^2, T, vT, dzT, vorticity^2 per row."""
        out = np.empty((self.ny, 9))
        self._c(self.lib.rbn_profiles(self.ctx, _dp(out)))
        return out

    def sample(self, field: str, x, y) -> np.ndarray:
        """Bilinear samples of rho/T/u/v of the latest step at positions (x, y)."""
        k = {"rho": 0, "T": 1, "u": 2, "v": 3}[field]
        x, y = _f64(x), _f64(y)
        out = np.empty(len(x))
        self._c(self.lib.rbn_sample(self.ctx, k, len(x), _dp(x), _dp(y), _dp(out)))
        return out

    # -- particles
    def init_particles(self, ids, x, y, tau_p=None, beta_p=None, up=None, vp=None) -> None:
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        n = len(ids)
        arrs = [_f64(a) for a in (x, y, tau_p, beta_p, up, vp)]
        self._c(self.lib.rbn_init_particles(self.ctx, n, ids.ctypes.data_as(C.POINTER(C.c_int)),
                                            *[_dp(a) for a in arrs]))

    def particles(self) -> dict:
        n = C.c_int()
        self._c(self.lib.rbn_num_particles(self.ctx, C.byref(n)))
        n = n.value
        out = {k: np.empty(max(n, 1)) for k in ("x", "y", "u", "v", "xo", "yo", "uo", "vo")}
        ids = np.empty(max(n, 1), dtype=np.int32)
        s = Particles()
        s.n = n
        s.id = ids.ctypes.data_as(C.POINTER(C.c_int))
        for k, a in out.items():
            setattr(s, k, _dp(a))
        self._c(self.lib.rbn_get_particles(self.ctx, C.byref(s)))
        res = {k: a[:n].copy() for k, a in out.items()}
        res["id"] = ids[:n].copy()
        return res

    def collect_records(self, fn: Optional[Callable[[dict], None]] = None) -> None:
        """Store every record (as numpy arrays) in self.records, or pass it to fn."""

        def cb(_user, recp):
            r = recp.contents
            n = r.n
            d = {"step": r.step, "id": np.ctypeslib.as_array(r.id, (n,)).copy() if n else np.empty(0, np.int32)}
            for k in RECORD_FIELDS + FULL_FIELDS:
                ptr = getattr(r, k)
                d[k] = np.ctypeslib.as_array(ptr, (n,)).copy() if (ptr and n) else None
            if fn is None:
                self.records.append(d)
            else:
                fn(d)
            return 0

        self._rec_cb = RECORD_FN(cb)
        self._c(self.lib.rbn_set_record_callback(self.ctx, self._rec_cb, None))

    def discard_records(self) -> None:
        """Deliver records to a callback that drops them (transfer cost only)."""
        self._rec_cb = RECORD_FN(lambda _u, _r: 0)
        self._c(self.lib.rbn_set_record_callback(self.ctx, self._rec_cb, None))

    def collect_diagnostics(self) -> None:
        def cb(_user, step, e, t):
            self.diagnostics.append((step, e, t))
            return 0

        self._diag_cb = DIAG_FN(cb)
        self._c(self.lib.rbn_set_diag_callback(self.ctx, self._diag_cb, None))

    # -- nudging
    def init_observers(self, ids, x, y, T) -> None:
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        arrs = [_f64(a) for a in (x, y, T)]
        self._c(self.lib.rbn_init_observers(self.ctx, len(ids), ids.ctypes.data_as(C.POINTER(C.c_int)),
                                            *[_dp(a) for a in arrs]))

    def set_observation_source(self, records: dict) -> None:
        """records: step -> dict(id, x, y, temp) (record storage order)."""

        def cb(_user, step, n, idp, xp, yp, tp):
            rec = records.get(step)
            if rec is None or len(rec["id"]) != n:
                return 1
            np.ctypeslib.as_array(idp, (n,))[:] = rec["id"]
            np.ctypeslib.as_array(xp, (n,))[:] = rec["x"]
            np.ctypeslib.as_array(yp, (n,))[:] = rec["y"]
            np.ctypeslib.as_array(tp, (n,))[:] = rec["temp"]
            return 0

        self._obs_cb = OBS_FN(cb)
        self._c(self.lib.rbn_set_obs_provider(self.ctx, self._obs_cb, None))

    def set_observation_table(self, records: dict) -> None:
        """As set_observation_source, served by the library from arrays kept alive here
        (every record must hold the same number of entries)."""
        steps = np.array(sorted(records), dtype=np.int32)
        n = len(records[int(steps[0])]["id"]) if len(steps) else 0
        cols = {}
        for key, dt in (("id", np.int32), ("x", np.float64), ("y", np.float64), ("temp", np.float64)):
            a = np.empty((len(steps), n), dtype=dt)
            for r, st in enumerate(steps):
                a[r] = records[int(st)][key]
            cols[key] = np.ascontiguousarray(a)
        self._obs_table = (steps, cols)
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
        self._c(self.lib.rbn_set_obs_table(self.ctx, len(steps), n, ip(steps), ip(cols["id"]), _dp(cols["x"]),
                                           _dp(cols["y"]), _dp(cols["temp"])))

    def nudging(self) -> dict:
        nv = np.empty((self.nx, self.ny), dtype=np.int32)
        tb = np.empty((self.nx, self.ny))
        nd = np.empty((self.nx, self.ny))
        self._c(self.lib.rbn_get_nudging(self.ctx, nv.ctypes.data_as(C.POINTER(C.c_int)), _dp(tb), _dp(nd)))
        w = np.empty((self.nx, self.ny))
        self._c(self.lib.rbn_get_nudging_weights(self.ctx, _dp(w)))
        return {"nudge_vol": nv, "tempreadint": tb, "nud_term": nd, "weight": w}

    # -- stepping
    def step(self, n: int = 1) -> None:
        self._c(self.lib.rbn_step(self.ctx, int(n)))

    @property
    def istep(self) -> int:
        s = C.c_int()
        self._c(self.lib.rbn_get_step(self.ctx, C.byref(s)))
        return s.value

    def synchronize(self) -> None:
        self._c(self.lib.rbn_synchronize(self.ctx))

    # -- adjoint window (nudged run)
    def adj_forward(self, nsteps: int, ckpt_every: int = 100) -> float:
        """Run nsteps nudged steps from the initial state and return the misfit J."""
        J = C.c_double()
        self._c(self.lib.rbn_adj_forward(self.ctx, int(nsteps), int(ckpt_every), C.byref(J)))
        return J.value

    def adj_backward(self) -> tuple[np.ndarray, float]:
        """Gradient of the last adj_forward's J: (dJ/dT0 as [nx][ny], dJ/dalpha)."""
        dT0 = np.empty((self.nx, self.ny))
        da = C.c_double()
        self._c(self.lib.rbn_adj_backward(self.ctx, _dp(dT0), C.byref(da)))
        return dT0, da.value

    def time_steps(self, n: int) -> float:
        ms = C.c_double()
        self._c(self.lib.rbn_time_steps(self.ctx, int(n), C.byref(ms)))
        return ms.value

    def set_tuning(self, **kw) -> None:
        """Launch shape of the fused step, e.g. set_tuning(lbm_block=512, lbm_nt=1, lbm_xcd=1)."""
        for k, v in kw.items():
            self._c(self.lib.rbn_set_tuning(self.ctx, k.encode(), int(v)))

    def tuning(self) -> dict:
        out = {}
        for k in ("lbm_block", "lbm_npt", "lbm_nt", "lbm_ntload", "lbm_uv_keep", "lbm_xcd", "lbm_reverse", "lbm_steps", "lbm3_seg", "lbm3_nt", "lbm2_seg", "lbm2_xcd", "lbm2_variant", "lbm2_tile_h", "particle_overlap", "particle_fuse", "dep_overlap",
                  "dep_in_tile", "obs_device", "adj_store_deposit"):
            v = C.c_int()
            self._c(self.lib.rbn_get_tuning(self.ctx, k.encode(), C.byref(v)))
            out[k] = v.value
        return out

    def kernel_timing(self, on, every: int = 1) -> None:
        """HIP-event timing of every `every`-th launch of each kernel class (off when on is false)."""
        self._c(self.lib.rbn_set_kernel_timing(self.ctx, int(every) if on else 0))

    def kernel_stats(self, name: str, reset: bool = False) -> tuple[int, float]:
        n, ms = C.c_longlong(), C.c_double()
        self._c(self.lib.rbn_kernel_stats(self.ctx, name.encode(), 1 if reset else 0, C.byref(n), C.byref(ms)))
        return n.value, ms.value


class Ensemble:
    """Several Solvers (truth-run or nudged) advanced together (rbn_ensemble_*)."""

    def __init__(self, members: list):
        self.lib = load()
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m.ctx for m in self.members])
        e = C.c_void_p()
        rc = self.lib.rbn_ensemble_create(arr, len(self.members), C.byref(e))
        if rc != 0:
            raise RbnError(rc, self.lib.rbn_last_error(None).decode())
        self.e = e

    def _c(self, rc: int) -> None:
        if rc != 0:
            raise RbnError(rc, self.lib.rbn_ensemble_last_error(self.e).decode())

    def close(self) -> None:
        if self.e:
            self.lib.rbn_ensemble_destroy(self.e)
            self.e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int) -> None:
        self._c(self.lib.rbn_ensemble_step(self.e, int(n)))

    def set_tuning(self, tile_w=256, tile_h=1, wg_per_cu=4, order=0) -> None:
        self._c(self.lib.rbn_ensemble_set_tuning(self.e, tile_w, tile_h, wg_per_cu, order))

    def tuning(self) -> dict:
        v = [C.c_int() for _ in range(4)]
        self._c(self.lib.rbn_ensemble_get_tuning(self.e, *[C.byref(x) for x in v]))
        return dict(zip(("tile_w", "tile_h", "wg_per_cu", "order"), (x.value for x in v)))

    def autotune(self, trial_steps: int = 8) -> float:
        ms = C.c_double()
        self._c(self.lib.rbn_ensemble_autotune(self.e, int(trial_steps), C.byref(ms)))
        return ms.value
