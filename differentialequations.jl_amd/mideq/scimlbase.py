"""Host-side mirror of the SciMLBase / OrdinaryDiffEq ensemble surface.

The reference path (DifferentialEquations.jl 8.0.3 re-exports these from
SciMLBase 3 and OrdinaryDiffEq 7; src/DifferentialEquations.jl:8-9):

    solve(EnsembleProblem(prob; prob_func, output_func, reduction),
          Tsit5() | Rodas5() | AutoTsit5(Rodas5()), EnsembleThreads();
          trajectories, saveat, abstol, reltol)

Names and argument meaning follow the reference (test/qa/qa.jl:15,50,52,56,
86,97,119,183,192,213; test/core.jl:13,31,46,84,93).  `EnsembleMI355X` is
the new EnsembleAlgorithm: its solve packs the problems that prob_func
returns into structure-of-arrays buffers and calls the C ABI of
libmideq.so (include/mideq.h), exactly as the Julia shim does with ccall.
Julia conventions kept: trajectory index i runs 1..N in prob_func and
output_func; output_func returns (out, rerun); reduction(u, batch, I)
returns (u, converged).
"""
import enum
import math
import time
from dataclasses import dataclass, field, replace
from typing import Any, Callable, Optional

import ctypes as C

import numpy as np

from . import _abi


# ---------------------------------------------------------------- return codes
class ReturnCode(enum.IntEnum):
    """SciMLBase.ReturnCode subset produced by the solver (qa.jl:213)."""
    Default = 0
    Success = 1
    MaxIters = 2
    DtLessThanMin = 3
    Unstable = 4
    DtNaN = 5
    InitialFailure = 6
    Terminated = 7


def successful_retcode(rc):
    return ReturnCode(int(rc)) == ReturnCode.Success


# ---------------------------------------------------------------- models
@dataclass(frozen=True)
class ODEFunction:
    """A right-hand side registered with the device library.

    `jac` says whether the model carries an analytic Jacobian (needed by
    Rodas5, as ODEFunction(f; jac=...) in the reference).  Calling the
    object evaluates f on the host (for inspection only; the solver never
    does)."""
    name: str
    n: Optional[int]
    np: int
    jac: bool = True

    @property
    def model_id(self):
        return _abi.MODEL[self.name]

    def __call__(self, u, p, t=0.0):
        u = np.asarray(u, dtype=np.float64)
        p = np.asarray(p, dtype=np.float64)
        if self.name == "exp":
            return p[0] * u
        if self.name == "lorenz":
            return np.array([p[0] * (u[1] - u[0]), u[0] * (p[1] - u[2]) - u[1],
                             u[0] * u[1] - p[2] * u[2]])
        if self.name == "robertson":
            return np.array([-p[0] * u[0] + p[2] * u[1] * u[2],
                             p[0] * u[0] - p[1] * u[1] ** 2 - p[2] * u[1] * u[2],
                             p[1] * u[1] ** 2])
        if self.name == "vdp":
            return np.array([u[1], p[0] * (1 - u[0] ** 2) * u[1] - u[0]])
        if self.name == "linear":
            return p.reshape(3, 3) @ u
        if self.name == "brusselator":
            N = len(u) // 2
            uu, vv = u[0::2], u[1::2]
            D = p[2] * N * N
            lu = np.roll(uu, 1) - 2 * uu + np.roll(uu, -1)
            lv = np.roll(vv, 1) - 2 * vv + np.roll(vv, -1)
            out = np.empty_like(u)
            out[0::2] = p[0] + uu * uu * vv - (p[1] + 1) * uu + D * lu
            out[1::2] = p[1] * uu - uu * uu * vv + D * lv
            return out
        raise ValueError(self.name)


@dataclass(frozen=True)
class ContinuousCallback:
    """ContinuousCallback(condition, affect!) (qa.jl:26) for the GPU:
    `condition` is a double expression of u[], p[], t whose sign change
    inside a step is the event; `affect` (statements changing u[]) runs at
    the located event time.  direction: 0 both crossings, 1 upcrossings,
    -1 downcrossings; terminate: end the trajectory (ReturnCode.Terminated)
    after the affect; interp_points: sign checks per step; save_positions:
    save (t, u) just before / just after the affect (upstream's default
    (True, True)), listed in the solution beside the saveat points."""
    condition: str
    affect: Optional[str] = None
    direction: int = 0
    terminate: bool = False
    interp_points: int = 10
    save_positions: tuple = (True, True)

    def spec(self):
        return _abi.CallbackSpec(0, int(self.direction), int(bool(self.terminate)),
                                 int(self.interp_points), self.condition.encode(),
                                 None if self.affect is None else self.affect.encode(),
                                 _save_bits(self.save_positions), 0)


@dataclass(frozen=True)
class DiscreteCallback:
    """DiscreteCallback(condition, affect!) (qa.jl:36) for the GPU: the
    affect runs at every step end where `condition` (an expression of u[],
    p[], t) is nonzero.  save_positions as for ContinuousCallback."""
    condition: str
    affect: Optional[str] = None
    terminate: bool = False
    save_positions: tuple = (True, True)

    def spec(self):
        return _abi.CallbackSpec(1, 0, int(bool(self.terminate)), 0, self.condition.encode(),
                                 None if self.affect is None else self.affect.encode(),
                                 _save_bits(self.save_positions), 0)


def _save_bits(save_positions):
    before, after = save_positions
    return (1 if before else 0) | (2 if after else 0)


@dataclass(frozen=True)
class DeviceOutputFunc:
    """output_func(sol, i) -> (out, false) of EnsembleProblem (qa.jl:50)
    evaluated on the GPU: statements assigning out[0..nout-1] from the final
    state u[], p[], the final time t and the retcode rc (compiled into the
    right-hand side).  Only the nout values per trajectory come back."""
    output: str
    nout: int = 1


@dataclass(frozen=True)
class DeviceReduction:
    """reduction(u, batch, I) of EnsembleProblem for the common folds,
    computed on the GPU over the DeviceOutputFunc values: "sum", "mean",
    "min" or "max" (fixed order: blocks of 256 by a halving tree, then block
    by block).  Only the nout reduced values come back."""
    op: str


def CallbackSet(*cbs):
    """CallbackSet (qa.jl:24): the callbacks as one tuple."""
    out = []
    for c in cbs:
        out += list(c) if isinstance(c, (list, tuple)) else [c]
    return tuple(out)


class UserModel:
    """A right-hand side given as C++ statements, compiled for the GPU at run
    time (hipRTC) -- the counterpart of passing an in-place f!(du, u, p, t)
    to ODEProblem (test/core.jl:22-27).  `rhs` assigns du[0..n-1] from u[i],
    p[j] and the time t (statements that name t make the model non-autonomous:
    stage times, and Rosenbrock's d_i dt df/dt terms with df/dt by dual
    numbers); its number type is T (double, float
    for FP32 stages, or a dual number when differentiated), so constants are
    written 2.0 or T(2), and sqrt exp log sin cos tanh fabs pow(x, c) are
    available.  `jac` assigns J[i][j] = df_i/du_j (double); without it the
    Rosenbrock methods differentiate rhs in forward mode.  A model with no
    parameters still takes one (unused) parameter row."""

    def __init__(self, n, np_, rhs, jac=None, name="user", callbacks=(), output=None,
                 save_everystep=False, with_tstops=False):
        if not (1 <= n <= _abi.USER_NMAX) or not (0 <= np_ <= _abi.USER_NPMAX):
            raise ValueError(f"user model n={n} np={np_} outside 1..{_abi.USER_NMAX} / "
                             f"0..{_abi.USER_NPMAX}")
        self.n = int(n)
        self.np = max(int(np_), 1)
        self.rhs = str(rhs)
        self.jac_src = None if jac is None else str(jac)
        self.jac = True
        self.name = name
        self.callbacks = CallbackSet(callbacks) if callbacks else ()
        if len(self.callbacks) > _abi.USER_NCBMAX:
            raise ValueError(f"at most {_abi.USER_NCBMAX} callbacks")
        self.output = output
        self.save_everystep = bool(save_everystep)
        self.with_tstops = bool(with_tstops)
        if output is not None and not (1 <= output.nout <= _abi.USER_NOUTMAX):
            raise ValueError(f"nout must be 1..{_abi.USER_NOUTMAX}")
        self._ids = {}
        self._with = {}

    def spec(self):
        cbs = [c.spec() for c in self.callbacks]
        arr = (_abi.CallbackSpec * max(1, len(cbs)))(*cbs)
        out = self.output
        sp = _abi.UserModelSpec(self.n, self.np, self.rhs.encode(),
                                None if self.jac_src is None else self.jac_src.encode(),
                                len(cbs), int(self.save_everystep), arr if cbs else None,
                                0 if out is None else int(out.nout), int(self.with_tstops),
                                None if out is None else out.output.encode())
        sp._keep = arr
        return sp

    def with_callbacks(self, callbacks=(), output=None, save_everystep=False,
                       with_tstops=False):
        """This model with `callbacks` (and an on-device output, save_everystep
        and tstops) compiled in (cached)."""
        cbs = CallbackSet(callbacks) if callbacks else ()
        save_everystep = bool(save_everystep) or self.save_everystep
        with_tstops = bool(with_tstops) or self.with_tstops
        key = (cbs, output, save_everystep, with_tstops)
        if key not in self._with:
            self._with[key] = UserModel(self.n, self.np, self.rhs, self.jac_src,
                                        name=self.name, callbacks=cbs, output=output,
                                        save_everystep=save_everystep,
                                        with_tstops=with_tstops)
        return self._with[key]

    def check(self):
        """Compile without a GPU; returns the compiler log (raises on error)."""
        L = _abi.load()
        buf = C.create_string_buffer(16384)
        sp = self.spec()
        rc = L.mideq_model_check(C.byref(sp), buf, len(buf))
        if rc != 0:
            raise _abi.MideqError(f"user model does not compile:\n{buf.value.decode()}")
        return buf.value.decode()

    def model_id_in(self, ctx):
        """The model id registered with `ctx` (registers on first use)."""
        key = id(ctx)
        if key not in self._ids:
            self._ids[key] = ctx.register_model(self.spec())
        return self._ids[key]

    def __call__(self, u, p, t=0.0):
        raise NotImplementedError("a user model runs on the GPU only")


# Device source of the built-in one-lane models, as user-model statements
# (the same arithmetic as their compiled form): callbacks compile into them.
BUILTIN_SOURCE = {
    "exp": (1, 1, "du[0] = p[0] * u[0];", "J[0][0] = p[0];"),
    "lorenz": (3, 3, "du[0] = p[0] * (u[1] - u[0]);\ndu[1] = u[0] * (p[1] - u[2]) - u[1];\n"
                     "du[2] = u[0] * u[1] - p[2] * u[2];",
               "J[0][0] = -p[0]; J[0][1] = p[0]; J[0][2] = 0.0;\n"
               "J[1][0] = p[1] - u[2]; J[1][1] = -1.0; J[1][2] = -u[0];\n"
               "J[2][0] = u[1]; J[2][1] = u[0]; J[2][2] = -p[2];"),
    "robertson": (3, 3, "T y1 = u[0], y2 = u[1], y3 = u[2];\nT k1 = p[0], k2 = p[1], k3 = p[2];\n"
                        "du[0] = -k1 * y1 + k3 * y2 * y3;\n"
                        "du[1] = k1 * y1 - k2 * (y2 * y2) - k3 * y2 * y3;\n"
                        "du[2] = k2 * (y2 * y2);",
                  "double y2 = u[1], y3 = u[2], k1 = p[0], k2 = p[1], k3 = p[2];\n"
                  "J[0][0] = -k1; J[0][1] = k3 * y3; J[0][2] = k3 * y2;\n"
                  "J[1][0] = k1; J[1][1] = -2.0 * k2 * y2 - k3 * y3; J[1][2] = -k3 * y2;\n"
                  "J[2][0] = 0.0; J[2][1] = 2.0 * k2 * y2; J[2][2] = 0.0;"),
    "vdp": (2, 1, "T x = u[0], y = u[1], mu = p[0];\ndu[0] = y;\n"
                  "du[1] = mu * (T(1) - x * x) * y - x;",
            "double x = u[0], y = u[1], mu = p[0];\nJ[0][0] = 0.0; J[0][1] = 1.0;\n"
            "J[1][0] = -2.0 * mu * x * y - 1.0; J[1][1] = mu * (1.0 - x * x);"),
}
_BUILTIN_USER = {}


def user_model_of(f):
    """f as a UserModel (built-in one-lane models through their source)."""
    if isinstance(f, UserModel):
        return f
    if f.name not in BUILTIN_SOURCE:
        raise ValueError(f"callbacks and save_everystep are not available for the {f.name} "
                         "model")
    if f.name not in _BUILTIN_USER:
        n, np_, rhs, jac = BUILTIN_SOURCE[f.name]
        _BUILTIN_USER[f.name] = UserModel(n, np_, rhs, jac=jac, name=f.name)
    return _BUILTIN_USER[f.name]


exp_growth = ODEFunction("exp", 1, 1)
lorenz = ODEFunction("lorenz", 3, 3)
robertson = ODEFunction("robertson", 3, 3)
vanderpol = ODEFunction("vdp", 2, 1)
linear3 = ODEFunction("linear", 3, 9)
brusselator = ODEFunction("brusselator", None, 3)


# ---------------------------------------------------------------- problems
@dataclass
class ODEProblem:
    """ODEProblem(f, u0, tspan, p) (qa.jl:86; test/core.jl:13,31,46)."""
    f: ODEFunction
    u0: Any
    tspan: tuple
    p: Any = None

    def __post_init__(self):
        self.u0 = np.atleast_1d(np.asarray(self.u0, dtype=np.float64))
        self.p = np.atleast_1d(np.asarray([] if self.p is None else self.p, dtype=np.float64))
        self.tspan = (float(self.tspan[0]), float(self.tspan[1]))


def remake(prob, u0=None, p=None, tspan=None, f=None):
    """remake(prob; u0, p, tspan) (qa.jl:183; test/core.jl:84)."""
    return ODEProblem(prob.f if f is None else f, prob.u0 if u0 is None else u0,
                      prob.tspan if tspan is None else tspan, prob.p if p is None else p)


def _default_prob_func(prob, i, repeat):
    return prob


def _default_output_func(sol, i):
    return sol, False


def _default_reduction(u, batch, I):
    return list(u) + list(batch), False


@dataclass
class EnsembleProblem:
    """EnsembleProblem(prob; prob_func, output_func, reduction, u_init,
    safetycopy) (qa.jl:50).  prob_func(prob, i, repeat) -> ODEProblem with
    i in 1..N.  A prob_func with attribute `vectorized = True` is instead
    called once as prob_func(prob, I, repeat) with the index array I and
    returns a dict(u0=[n,N], p=[np,N], t0=..., tf=...), which avoids N
    Python calls for large ensembles."""
    prob: ODEProblem
    prob_func: Callable = _default_prob_func
    output_func: Callable = _default_output_func
    reduction: Callable = _default_reduction
    u_init: Any = None
    safetycopy: bool = True


# ---------------------------------------------------------------- algorithms
@dataclass(frozen=True)
class Tsit5:
    """Tsitouras 5(4) (qa.jl:119)."""
    code = _abi.ALG["tsit5"]


@dataclass(frozen=True)
class Rodas5:
    """8-stage stiffly accurate Rosenbrock with analytic Jacobian."""
    code = _abi.ALG["rodas5"]


@dataclass(frozen=True)
class Rodas5P:
    """Rodas5P (qa.jl:97): Rodas5's stage structure with Steinebach's 2023
    coefficients, the Rosenbrock that DifferentialEquations re-exports."""
    code = _abi.ALG["rodas5p"]


@dataclass(frozen=True)
class AutoTsit5:
    """AutoTsit5(Rodas5()) / AutoTsit5(Rodas5P()) (qa.jl:15): Tsit5 with
    stiffness detection, switching to the given Rosenbrock method."""
    stiffalg: Any = field(default_factory=Rodas5)

    def __post_init__(self):
        if not isinstance(self.stiffalg, (Rodas5, Rodas5P)):
            raise TypeError(f"AutoTsit5 supports Rodas5() and Rodas5P(), got {self.stiffalg!r}")

    @property
    def code(self):
        return _abi.ALG["auto5p"] if isinstance(self.stiffalg, Rodas5P) else _abi.ALG["auto"]


@dataclass(frozen=True)
class EnsembleMI355X:
    """EnsembleAlgorithm that solves every trajectory on one MI355X.

    precision: "fp64" or "mixed" (FP32 stages; FP64 state update, norm,
    time, J/W/LU).  schedule: "auto", "lane" or "persistent".
    dt0 / alg_hint: optional per-trajectory initial step and initial method
    (e.g. from mideq.learned).  order: optional launch order, a permutation
    of the trajectory indices (0-based) -- consecutive entries share a
    wavefront, so grouping trajectories of similar cost cuts lane
    divergence; results keep their natural positions.  device: GPU ordinal.
    auto_passes (AutoTsit5 on n <= 3 models): 0/1 = one AutoSwitch pass;
    2 = a Tsit5 pass, then the full AutoSwitch kernel over the trajectories
    that switch (faster when few switch).  Both give identical results.
    3 = trust alg_hint: plain Tsit5 for the trajectories hinted non-stiff,
    each capped at iter_cap[i] attempted steps, then AutoSwitch from t0 for
    the capped-out and the hinted-stiff ones.  save_layout: "soa" (saved
    values [ns, n, N]) or "trajectory" ([N, ns, n], one trajectory's saved
    states contiguous, as upstream's per-trajectory sol.u).  resident: the
    packed problems go to the GPU once and every output stays there as a
    torch tensor; each ODESolution's states are views of those tensors (its
    times and counts are read back once per solve)."""
    precision: str = "fp64"
    schedule: str = "auto"
    device: int = 0
    save_dtype: str = "float64"
    save_layout: str = "soa"
    dt0: Any = None
    alg_hint: Any = None
    order: Any = None
    lanes: int = 0
    graph: bool = False
    controller: str = "reference"   # or "fast" (n <= 3 models)
    auto_passes: int = 0
    iter_cap: Any = None
    resident: bool = False


# ---------------------------------------------------------------- solutions
@dataclass
class Stats:
    naccept: int
    nreject: int
    nf: int


@dataclass
class ODESolution:
    """One trajectory's result (qa.jl:87)."""
    t: np.ndarray
    u: list
    retcode: ReturnCode
    stats: Stats
    prob: Optional[ODEProblem] = None
    interp: Optional[Callable] = field(default=None, repr=False, compare=False)

    def __getitem__(self, k):
        return self.u[k]

    def __call__(self, t):
        """Dense output sol(t) (core.jl:51-58): the state at time t (a
        scalar, or an array of times -> [len(t), n]) from the interpolant of
        the step that covers t; NaN outside the integrated span."""
        if self.interp is None:
            raise ValueError("this solution has no dense output: solve with dense=True "
                             "(the default when no saveat is given)")
        vals = self.interp(np.atleast_1d(np.asarray(t, dtype=np.float64)))
        return vals[0] if np.ndim(t) == 0 else vals


@dataclass
class EnsembleArrays:
    """Structure-of-arrays result of one device solve: numpy arrays, or for
    a device-resident solve (resident = True) torch tensors on the GPU whose
    per-trajectory views saved(i), events(i) and solution(i) copy no state
    (the zero-copy EnsembleSolution view of qa.jl:52)."""
    u_final: np.ndarray           # [n, N]
    t_final: np.ndarray           # [N]
    retcode: np.ndarray           # [N] int8
    naccept: np.ndarray
    nreject: np.ndarray
    nf: np.ndarray
    nstiff: np.ndarray
    dt_first: np.ndarray          # [N] first accepted step
    saveat: Optional[np.ndarray]  # [ns]
    us: Optional[np.ndarray]      # [ns, n, N], or [N, ns, n] (save_layout "trajectory")
    t0: np.ndarray
    u0: np.ndarray
    kernel_ms: float = 0.0
    total_ms: float = 0.0
    lanes: int = 0
    graph: bool = False
    grid: int = 0                 # blocks launched (persistent kernels: resident teams' blocks)
    nevent: Optional[np.ndarray] = None   # [N] callback affects applied
    t_event: Optional[np.ndarray] = None  # [N] time of the last one (NaN if none)
    out: Optional[np.ndarray] = None      # [nout, N] on-device output_func values
    reduced: Optional[np.ndarray] = None  # [nout] their on-device reduction
    save_layout: str = "soa"
    nvec: Optional[np.ndarray] = None     # [N] state sizes of a ragged batch
    ev_t: Optional[np.ndarray] = None     # [N, cap] times of the points saved at events
    ev_u: Optional[np.ndarray] = None     # [N, cap, n] their states
    ev_n: Optional[np.ndarray] = None     # [N] points produced (may exceed cap)
    resident: bool = False                # the arrays are torch tensors on the GPU
    save_everystep: bool = False          # ev_* also hold every accepted step's end point
    drop_start: bool = False              # save_start = false: no (t0, u0) point
    drop_end: bool = False                # save_end = false: no final point appended
    keep: Optional[list] = field(default=None, repr=False, compare=False)  # buffers in use
    _meta: Optional[dict] = field(default=None, repr=False, compare=False)

    def _h(self, name):
        """Per-trajectory scalars (counts, codes, times) as host arrays: for a
        device-resident result read back once, on first use (the states stay
        on the GPU)."""
        if not self.resident:
            return getattr(self, name)
        if self._meta is None:
            names = ("naccept", "nreject", "nf", "retcode", "t_final", "t0", "nvec", "ev_n")
            self._meta = {k: (None if getattr(self, k) is None else getattr(self, k).cpu().numpy())
                          for k in names}
        return self._meta[name]

    def saved(self, i):
        """Saved states of trajectory i (0-based), [ns, n_i]."""
        vals = self.us[i] if self.save_layout == "trajectory" else self.us[:, :, i]
        return vals[:, :int(self._h("nvec")[i])] if self.nvec is not None else vals

    def events(self, i):
        """(times, states) of the points trajectory i saved at callback events,
        in the order of integration (at most the arrays' capacity).  For a
        device-resident result the times are read back and the states are a
        view on the GPU."""
        if self.ev_t is None:
            return np.empty(0), np.empty((0, self.u0.shape[0]))
        m = min(int(self._h("ev_n")[i]), self.ev_t.shape[1])
        if self.resident:
            return self.ev_t[i, :m].cpu().numpy(), self.ev_u[i, :m]
        return self.ev_t[i, :m].copy(), self.ev_u[i, :m].copy()

    def solution(self, i, prob=None):
        """ODESolution of trajectory i (0-based).  Points saved at callback
        events and, with save_everystep, at every step are listed with the
        saveat points in the order of integration (a saveat point first at an
        equal time, as upstream's savevalues! saves it first); (t0, u0)
        starts the solution without saveat or with save_everystep, and
        without saveat the final state ends it unless the last point saved
        already has its time (upstream's postamble)."""
        h = self._h
        st = Stats(int(h("naccept")[i]), int(h("nreject")[i]), int(h("nf")[i]))
        rc = ReturnCode(int(h("retcode")[i]))
        back = self.t_final is not None and float(h("t_final")[i]) < float(h("t0")[i])
        et, eu = self.events(i)
        # host results hand out copies, device-resident ones views
        own = (lambda v: v) if self.resident else np.array
        if self.us is not None:
            vals = self.saved(i)
            nan = vals.isnan().any(dim=1).cpu().numpy() if self.resident else \
                np.isnan(vals).any(axis=1)
            rows = np.nonzero(~nan)[0]
            ts, us = self.saveat[rows], [vals[k] for k in rows]
            if back:
                ts, us = ts[::-1], us[::-1]  # a backward solve: in the order of integration
            if len(et):
                allt = np.concatenate([ts, et])
                allu = list(us) + [eu[k] for k in range(len(et))]
                order = np.argsort(-allt if back else allt, kind="stable")
                ts, us = allt[order], [allu[k] for k in order]
            t0 = float(h("t0")[i])
            if self.save_everystep and (len(ts) == 0 or ts[0] != t0):
                m = int(h("nvec")[i]) if self.nvec is not None else self.u0.shape[0]
                ts, us = np.concatenate([[t0], ts]), [self.u0[:m, i]] + list(us)
            return ODESolution(np.array(ts, dtype=np.float64), [own(v) for v in us], rc, st,
                               prob)
        m = int(h("nvec")[i]) if self.nvec is not None else self.u0.shape[0]
        tf = float(h("t_final")[i])
        ts = np.concatenate([[float(h("t0")[i])], et])
        us = [own(self.u0[:m, i])] + [eu[k, :m] for k in range(len(et))]
        cut = self.ev_t is not None and int(self._h("ev_n")[i]) > self.ev_t.shape[1]
        if (len(et) == 0 or et[-1] != tf or cut) and not self.drop_end:
            ts, us = np.concatenate([ts, [tf]]), us + [own(self.u_final[:m, i])]
        if self.drop_start:
            ts, us = ts[1:], us[1:]
        return ODESolution(ts, us, rc, st, prob)


@dataclass
class EnsembleSolution:
    """EnsembleSolution (qa.jl:52): outputs in trajectory order."""
    u: Any
    elapsedTime: float
    converged: bool
    arrays: Optional[EnsembleArrays] = None

    def __len__(self):
        return len(self.u)

    def __getitem__(self, k):
        return self.u[k]


# ---------------------------------------------------------------- packing
def _pack(ens, idx, repeat):
    """Evaluate prob_func for trajectories idx (0-based) and pack SoA."""
    pf = ens.prob_func
    if getattr(pf, "vectorized", False):
        d = pf(ens.prob, idx + 1, repeat)
        u0 = np.ascontiguousarray(d["u0"], dtype=np.float64)
        p = np.ascontiguousarray(d["p"], dtype=np.float64)
        N = len(idx)
        t0 = np.ascontiguousarray(np.broadcast_to(d.get("t0", ens.prob.tspan[0]), (N,)),
                                  dtype=np.float64)
        tf = np.ascontiguousarray(np.broadcast_to(d.get("tf", ens.prob.tspan[1]), (N,)),
                                  dtype=np.float64)
        nvec = d.get("nvec")
        return ens.prob.f, u0, p, t0, tf, nvec, None
    probs = []
    for i in idx:
        base = ens.prob
        pr = pf(base, int(i) + 1, repeat)
        probs.append(pr)
    f = probs[0].f
    for pr in probs:
        if pr.f is not f and pr.f != f:
            raise ValueError("all trajectories of one ensemble must share the RHS")
    if f.name == "brusselator":
        nmax = max(len(pr.u0) for pr in probs)
        u0 = np.zeros((nmax, len(probs)))
        for k, pr in enumerate(probs):
            u0[:len(pr.u0), k] = pr.u0
        nvec = np.array([len(pr.u0) for pr in probs], dtype=np.int32)
    else:
        u0 = np.ascontiguousarray(np.stack([pr.u0 for pr in probs], axis=1))
        nvec = None
    p = np.ascontiguousarray(np.stack([pr.p for pr in probs], axis=1))
    t0 = np.array([pr.tspan[0] for pr in probs])
    tf = np.array([pr.tspan[1] for pr in probs])
    return f, u0, p, t0, tf, nvec, probs


def _rat(x):
    """Continued-fraction rational approximation of a float, as Julia's
    Base.rat does for Float64 ranges (coefficients bounded by
    maxintfloat(Float32) = 2^24): returns (num, den), den == 0 if none."""
    m = 16777216.0
    y = x
    a, d = 1, 1
    b, c = 0, 0
    while abs(y) <= m:
        f = int(y)  # trunc
        y -= f
        a, c = f * a + c, a
        b, d = f * b + d, b
        if max(abs(a), abs(b)) > m:
            return c, d
        if b != 0 and float(a) / float(b) == x:
            break
        if y == 0.0:
            break
        y = 1.0 / y
    return a, b


def julia_range(start, step, stop):
    """The elements of Julia's `start:step:stop` for Float64 (base/
    twiceprecision.jl semantics): when start, step and stop are exact
    rationals p/q (found by `_rat`), element k is the correctly rounded
    value of (start_n + k step_n) / den, so 0:0.1:0.3 ends at 0.3 exactly;
    otherwise element k is start + k step correctly rounded, and the length
    is round((stop - start) / step) + 1, minus one on overshoot."""
    from fractions import Fraction
    start, step, stop = float(start), float(step), float(stop)
    if step == 0.0:
        raise ValueError("range step cannot be zero")

    def between(a, x, b):
        return (a <= x <= b) or (b <= x <= a)

    def tdiv(p, q):  # Julia's div: truncation toward zero
        r = abs(p) // abs(q)
        return r if (p >= 0) == (q > 0) else -r

    sn, sd = _rat(step)
    if sd != 0 and sn / sd == step:
        an, ad = _rat(start)
        zn, zd = _rat(stop)
        if ad != 0 and zd != 0 and an / ad == start and zn / zd == stop:
            den = (ad // math.gcd(ad, sd)) * sd  # Julia's lcm_unchecked (sign kept)
            m = 2.0 ** 53
            if abs(start * den) <= m and abs(step * den) <= m:
                a_n = int(round(start * den))
                s_n = int(round(step * den))
                ln = max(0, tdiv(den * zn - zd * a_n + s_n * zd, s_n * zd))
                if (between(start, start + (ln - 1) * step, stop + step / 2)
                        and not between(start, start + ln * step, stop)):
                    return np.array([float(Fraction(a_n + k * s_n, den)) for k in range(ln)])
    lf = (stop - start) / step
    if lf < 0:
        ln = 0
    elif lf == 0:
        ln = 1
    else:
        ln = int(round(lf)) + 1
        stop2 = start + (ln - 1) * step
        ln -= int(start < stop < stop2) + int(start > stop > stop2)
    fs, fd = Fraction(start), Fraction(step)
    return np.array([float(fs + k * fd) for k in range(ln)])


def _saveat_vector(saveat, tspan, save_start=None, save_end=None):
    """The saveat points handed to the kernels, ascending.  save_start /
    save_end (None: upstream's defaults -- true for a scalar saveat, else
    whether t0 / tf is among the points) decide whether t0 and tf are saved,
    as OrdinaryDiffEq keeps only the points strictly inside the span in its
    queue and saves the ends by these flags."""
    if saveat is None:
        return None
    t0, t1 = float(tspan[0]), float(tspan[1])
    if np.isscalar(saveat):
        # the range t0:dir*d:tf, then tf itself when the range stops short of
        # it (save_end), in the direction of integration; handed over in
        # ascending order (a backward solve fills them from the end)
        sgn = 1.0 if t1 >= t0 else -1.0
        pts = julia_range(t0, sgn * abs(float(saveat)), t1)
        if len(pts) == 0 or pts[-1] != t1:
            pts = np.append(pts, t1)
    else:
        pts = np.asarray(saveat, dtype=np.float64).ravel()
    for end, flag in ((t0, save_start), (t1, save_end)):
        if flag is not None:
            pts = pts[pts != end]
            if flag:
                pts = np.append(pts, end)
    return np.ascontiguousarray(np.sort(pts), dtype=np.float64)


def _ptr(a):
    if a is None:
        return None
    return a.data_ptr() if hasattr(a, "data_ptr") else a.ctypes.data


class _HostMem:
    """Buffers of a host-entry solve: numpy arrays."""
    resident = False

    def arr(self, a, dtype):
        if hasattr(a, "detach"):  # a torch tensor in host memory
            a = a.detach().numpy()
        return np.ascontiguousarray(a, dtype=dtype)

    def full(self, shape, value, dtype):
        return np.full(shape, value, dtype=dtype)

    def empty(self, shape, dtype):
        return np.empty(shape, dtype=dtype)


class _DeviceMem:
    """Buffers of a device-resident solve: torch tensors on one GPU.  Inputs
    already there (contiguous, of the stated dtype) are used in place, others
    are copied there once; outputs are allocated there and stay there.
    Every buffer is kept (and recorded on the solve's stream) until the
    result that holds it is dropped, so the asynchronous solve never sees
    memory handed to another allocation."""
    resident = True

    def __init__(self, torch, device):
        self.torch, self.device, self.keep = torch, device, []
        self.dtypes = {np.dtype(np.float64): torch.float64, np.dtype(np.float32): torch.float32,
                       np.dtype(np.int64): torch.int64, np.dtype(np.int32): torch.int32,
                       np.dtype(np.int8): torch.int8}

    def _k(self, t):
        self.keep.append(t)
        return t

    def arr(self, a, dtype):
        T = self.torch
        if not T.is_tensor(a):
            a = T.from_numpy(np.ascontiguousarray(a, dtype=dtype))
        return self._k(a.to(device=self.device, dtype=self.dtypes[np.dtype(dtype)]).contiguous())

    def full(self, shape, value, dtype):
        return self._k(self.torch.full(tuple(np.atleast_1d(shape).tolist()), value,
                                       dtype=self.dtypes[np.dtype(dtype)], device=self.device))

    def empty(self, shape, dtype):
        return self._k(self.torch.empty(tuple(np.atleast_1d(shape).tolist()),
                                        dtype=self.dtypes[np.dtype(dtype)], device=self.device))


def _resident_device(u0):
    """(torch, device) when u0 is a torch tensor in GPU memory, else None."""
    if not (hasattr(u0, "is_cuda") and u0.is_cuda):
        return None
    import torch
    return torch, u0.device


def device_solve(f, u0, p, t0, tf, nvec, alg, ealg, saveat=None, abstol=1e-6, reltol=1e-3,
                 dt=0.0, dtmin=0.0, dtmax=0.0, maxiters=0, dt0=None, alg_hint=None,
                 order=None, block=0, blocks_per_cu=0, callback=None, output=None,
                 reduction=None, keep_states=True, event_cap=16, stream=None,
                 save_everystep=False, tstops=None):
    """One call of mideq_ensemble_solve on host arrays -> EnsembleArrays.
    callback: ContinuousCallback / DiscreteCallback / CallbackSet, compiled
    into the right-hand side (a UserModel, or a built-in model's source).
    output: DeviceOutputFunc (values per trajectory in .out); reduction:
    DeviceReduction or its op name (folded values in .reduced, and with
    keep_states=False nothing per trajectory comes back).  save_everystep:
    every accepted step's end point is kept too (one-lane models; built-in
    ones through their source), with the event points, event_cap points per
    trajectory at most (the count goes on past it).  tstops: times every
    trajectory's steps land on (upstream's kwarg; one-lane models).

    Device-resident form: when u0 is a torch tensor in GPU memory the call
    goes to mideq_ensemble_solve_device instead.  The other inputs are used
    in place when they are already there (else copied there once), every
    output is a torch tensor on that GPU, nothing comes back to the host, and
    the call returns once the solve is enqueued on `stream` (a
    torch.cuda.Stream; default the current stream) -- synchronize before
    reading.  The result's saved(i) and solution(i) are views into those
    tensors (no copies)."""
    if tstops is not None:
        tstops = np.sort(np.asarray(tstops, dtype=np.float64).ravel())
        if len(tstops) == 0:
            tstops = None
    if callback or output is not None or save_everystep or tstops is not None:
        f = user_model_of(f).with_callbacks(callback or (), output, save_everystep,
                                            tstops is not None)
    if isinstance(reduction, DeviceReduction):
        reduction = reduction.op
    dev = _resident_device(u0)
    mem = _DeviceMem(*dev) if dev else _HostMem()
    if dev is not None:
        torch, device = dev
        if ealg.device != device.index:
            raise ValueError(f"u0 is on {device}, EnsembleMI355X(device={ealg.device})")
        if stream is None:
            stream = torch.cuda.current_stream(device)
    elif stream is not None:
        raise ValueError("stream is for device-resident solves (u0 a GPU tensor)")
    ctx = _abi.context(ealg.device)
    n, N = u0.shape
    u0 = mem.arr(u0, np.float64)
    p = mem.arr(p, np.float64)
    P = _abi.Problem()
    P.model = f.model_id_in(ctx) if isinstance(f, UserModel) else f.model_id
    P.n = n
    P.np = p.shape[0]
    P.ntraj = N
    P.u0 = _ptr(u0)
    P.p = _ptr(p)
    if not mem.resident:
        P.tspan0 = float(np.asarray(t0).ravel()[0]) if N else 0.0
        P.tspan1 = float(np.asarray(tf).ravel()[0]) if N else 0.0
    t0 = mem.arr(t0, np.float64)
    tf = mem.arr(tf, np.float64)
    P.t0 = _ptr(t0)
    P.tf = _ptr(tf)
    if nvec is not None:
        nvec = mem.arr(nvec, np.int32)
        P.nvec = _ptr(nvec)
    O = _abi.Options()
    O.alg = alg.code
    O.precision = _abi.PRECISION[ealg.precision]
    O.abstol = abstol
    O.reltol = reltol
    O.dt = dt
    O.dtmin = dtmin
    O.dtmax = dtmax
    O.maxiters = maxiters
    O.qsteady_default = 1
    O.schedule = _abi.SCHEDULE[ealg.schedule]
    O.block = block
    O.blocks_per_cu = blocks_per_cu
    O.lanes = ealg.lanes
    O.graph = 1 if ealg.graph else 0
    O.controller = dict(reference=0, fast=1)[ealg.controller]
    O.auto_passes = int(ealg.auto_passes)
    iter_cap = None
    if ealg.iter_cap is not None:
        iter_cap = mem.arr(ealg.iter_cap, np.int32)
        if iter_cap.shape != (N,):
            raise ValueError(f"iter_cap has shape {iter_cap.shape}, expected ({N},)")
        O.iter_cap = _ptr(iter_cap)
    f32 = ealg.save_dtype == "float32"
    O.save_dtype = 1 if f32 else 0
    us = None
    if saveat is not None and len(saveat):
        # the times as given (host) for the solution; a copy where the solve reads
        saveat_in = saveat
        saveat = np.ascontiguousarray(saveat.cpu().numpy() if hasattr(saveat, "cpu")
                                      else saveat, dtype=np.float64)
        O.saveat = _ptr(mem.arr(saveat_in, np.float64) if mem.resident else saveat)
        O.ns = len(saveat)
        O.save_layout = dict(soa=0, trajectory=1)[ealg.save_layout]
        shape = (N, len(saveat), n) if ealg.save_layout == "trajectory" else (len(saveat), n, N)
        # rows past a ragged instance's size are left as they are: NaN
        us = mem.full(shape, np.nan, np.float32 if f32 else np.float64)
    dt0 = dt0 if dt0 is not None else ealg.dt0
    hint = alg_hint if alg_hint is not None else ealg.alg_hint
    if dt0 is not None:
        dt0 = mem.arr(dt0, np.float64)
        O.dt0 = _ptr(dt0)
    if hint is not None:
        hint = mem.arr(hint, np.int8)
        O.alg_hint = _ptr(hint)
    order = order if order is not None else ealg.order
    if order is not None:
        order = mem.arr(order, np.int64)
        if order.shape != (N,):
            raise ValueError(f"order has shape {order.shape}, expected ({N},)")
        O.order = _ptr(order)
    if reduction is not None:
        O.reduction = _abi.REDUCTION[reduction]
    if tstops is not None:
        tstops = mem.arr(tstops, np.float64)
        O.tstops = _ptr(tstops)
        O.ntstops = len(tstops)
    if keep_states:
        res = dict(u_final=mem.full((n, N), 0.0, np.float64), t_final=mem.empty(N, np.float64),
                   retcode=mem.empty(N, np.int8), naccept=mem.empty(N, np.int32),
                   nreject=mem.empty(N, np.int32), nf=mem.empty(N, np.int32),
                   nstiff=mem.empty(N, np.int32), dt_first=mem.empty(N, np.float64))
    else:
        res = dict(u_final=None, t_final=None, retcode=None, naccept=None, nreject=None,
                   nf=None, nstiff=None, dt_first=None)
    everystep = isinstance(f, UserModel) and f.save_everystep
    if isinstance(f, UserModel) and (f.callbacks or everystep) and keep_states:
        if f.callbacks:
            res["nevent"] = mem.full(N, 0, np.int32)
            res["t_event"] = mem.empty(N, np.float64)
        if everystep or any(_save_bits(getattr(cb, "save_positions", (False, False)))
                            for cb in f.callbacks):
            # points saved at events (save_positions): event_cap per trajectory
            res["ev_t"] = mem.full((N, int(event_cap)), np.nan, np.float64)
            res["ev_u"] = mem.full((N, int(event_cap), n), np.nan, np.float64)
            res["ev_n"] = mem.full(N, 0, np.int32)
    nout = f.output.nout if isinstance(f, UserModel) and f.output is not None else 0
    if nout and (keep_states or (mem.resident and reduction is not None)):
        # the device entry point folds the per-trajectory values from a
        # buffer of the caller's (the host entry point makes its own)
        res["out"] = mem.empty((nout, N), np.float64)
    if nout and reduction is not None:
        res["reduced"] = mem.empty(nout, np.float64)
    out = _abi.Outputs()
    out.us = _ptr(us)
    for k, v in res.items():
        setattr(out, k, _ptr(v))
    if "ev_t" in res:
        out.ev_cap = int(event_cap)
    if mem.resident:
        cur = torch.cuda.current_stream(device)
        if stream != cur:  # the buffers above were filled on the current stream
            stream.wait_stream(cur)
            for t in mem.keep:
                t.record_stream(stream)
        ctx.solve_device(P, O, out, C.c_void_p(stream.cuda_stream))
    else:
        ctx.solve(P, O, out)
    st = ctx.stats()
    return EnsembleArrays(saveat=saveat, us=us, t0=t0, u0=u0, kernel_ms=st.kernel_ms,
                          total_ms=st.total_ms, lanes=st.lanes, grid=st.grid,
                          graph=bool(st.graph),
                          save_layout=ealg.save_layout, nvec=nvec,
                          save_everystep=bool(everystep and keep_states),
                          resident=mem.resident, keep=mem.keep if mem.resident else None,
                          **res)


def _dense_interp(packed, k, i, alg, ealg, lanes, opts, kw):
    """sol(t) for trajectory i (0-based), column k of the packed inputs
    (f, u0, p, t0, tf, nvec): the same solve again with the requested times
    as its saveat points.  Saved points do not change the steps taken (there
    are no tstops at them), so the values are the interpolants of the steps
    of the solve that made the solution, bit for bit.  lanes: the team shape
    that solve used (the Brusselator team's reduction order depends on it).
    Nothing is copied until the solution is called."""

    def interp(ts):
        f, u0, p, t0, tf, nvec = packed
        one = np.array([i])
        ealg1 = replace(_subset_ealg(ealg, one), lanes=lanes, graph=False, save_layout="soa",
                        save_dtype="float64", order=None)
        extra = {key: np.asarray(kw[key])[[k]] for key in ("dt0", "alg_hint")
                 if kw.get(key) is not None}
        for key in ("callback", "tstops"):  # the same steps: the same events and tstops
            if kw.get(key) is not None:
                extra[key] = kw[key]
        ts = np.ascontiguousarray(ts, dtype=np.float64).ravel()
        order = np.argsort(ts, kind="stable")
        sa = np.ascontiguousarray(ts[order])
        arr = device_solve(f, np.ascontiguousarray(u0[:, k:k + 1]),
                           np.ascontiguousarray(p[:, k:k + 1]), np.array(t0[k:k + 1]),
                           np.array(tf[k:k + 1]),
                           None if nvec is None else np.asarray(nvec)[k:k + 1].copy(),
                           alg, ealg1, saveat=sa, **opts, **extra)
        vals = arr.saved(0)
        out = np.empty_like(vals)
        out[order] = vals
        return out

    return interp


def _resident_inputs(ealg, u0, p, t0, tf, nvec):
    """The packed inputs on the GPU (torch tensors) for a resident solve."""
    if not ealg.resident:
        return u0, p, t0, tf, nvec
    import torch
    dev = torch.device("cuda", ealg.device)

    def put(a, dt):
        return None if a is None else torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)

    return (put(u0, np.float64), put(p, np.float64), put(t0, np.float64), put(tf, np.float64),
            put(nvec, np.int32))


def _subset_ealg(ealg, pending):
    """The per-trajectory settings of `ealg` restricted to the trajectories
    `pending` (a rerun pass), the launch order kept as their relative order."""
    kw = {}
    if ealg.dt0 is not None:
        kw["dt0"] = np.asarray(ealg.dt0)[pending]
    if ealg.alg_hint is not None:
        kw["alg_hint"] = np.asarray(ealg.alg_hint)[pending]
    if ealg.iter_cap is not None:
        kw["iter_cap"] = np.asarray(ealg.iter_cap)[pending]
    if ealg.order is not None:
        rank = np.empty(len(ealg.order), dtype=np.int64)
        rank[np.asarray(ealg.order, dtype=np.int64)] = np.arange(len(ealg.order))
        kw["order"] = np.argsort(rank[pending], kind="stable")
    return replace(ealg, **kw) if kw else ealg


# ---------------------------------------------------------------- solve
def solve(prob, alg, ensemblealg=None, *, trajectories=None, saveat=None, abstol=1e-6,
          reltol=1e-3, dt=0.0, dtmin=0.0, dtmax=0.0, maxiters=100000, **kw):
    """solve(EnsembleProblem, alg, EnsembleMI355X(); trajectories, ...)
    (qa.jl:192) or solve(ODEProblem, alg) for a single trajectory.

    save_start / save_end (upstream's flags): whether t0 and tf are saved
    (defaults as upstream: with a scalar saveat or none, both; with a vector,
    whether it holds them).

    save_everystep=True keeps every accepted step's end point in each
    solution (upstream's default without saveat; here opt-in, since it
    compiles the model with the step saves; sol(t) interpolates without it).
    The points per trajectory are capped by event_cap, by default what
    ~1 GiB holds over the ensemble (at least 16, at most 65,536); past the
    cap a solution ends with the final state after the points kept."""
    if isinstance(prob, ODEProblem):
        ens = EnsembleProblem(prob)
        es = solve(ens, alg, ensemblealg or EnsembleMI355X(), trajectories=1, saveat=saveat,
                   abstol=abstol, reltol=reltol, dt=dt, dtmin=dtmin, dtmax=dtmax,
                   maxiters=maxiters, **kw)
        return es.u[0]
    if not isinstance(prob, EnsembleProblem):
        raise TypeError("solve expects an ODEProblem or EnsembleProblem")
    if ensemblealg is None:
        ensemblealg = EnsembleMI355X()
    if not isinstance(ensemblealg, EnsembleMI355X):
        raise TypeError(f"unsupported ensemble algorithm {ensemblealg!r}")
    if trajectories is None:
        raise TypeError("trajectories is required")
    tstart = time.perf_counter()
    N = int(trajectories)
    if kw.get("save_everystep") and "event_cap" not in kw:
        n0 = max(1, int(np.size(prob.prob.u0)))
        kw["event_cap"] = int(min(max(16, (1 << 27) // (max(N, 1) * (n0 + 1))),
                                  min(int(maxiters), 65536)))
    batch_size = int(kw.pop("batch_size", N))
    if batch_size < 1 or N // batch_size < 1:
        raise ValueError("trajectories // batch_size cannot be less than 1")
    idx = np.arange(N)
    save_start, save_end = kw.pop("save_start", None), kw.pop("save_end", None)
    sa = _saveat_vector(saveat, prob.prob.tspan, save_start, save_end)
    # dense output (sol(t)): on by default without saveat, as upstream
    dense = kw.pop("dense", None)
    dense = saveat is None if dense is None else bool(dense)
    if isinstance(prob.output_func, DeviceOutputFunc):
        # output_func (and reduction) on the device: one pass, no ODESolutions
        red = prob.reduction if isinstance(prob.reduction, DeviceReduction) else None
        f, u0, p, t0, tf, nvec, probs = _pack(prob, idx, 1)
        u0, p, t0, tf, nvec = _resident_inputs(ensemblealg, u0, p, t0, tf, nvec)
        arr = device_solve(f, u0, p, t0, tf, nvec, alg, ensemblealg, saveat=None,
                           abstol=abstol, reltol=reltol, dt=dt, dtmin=dtmin, dtmax=dtmax,
                           maxiters=maxiters, output=prob.output_func, reduction=red,
                           keep_states=red is None, **kw)
        u = arr.reduced if red is not None else arr.out.T
        return EnsembleSolution(u=u, elapsedTime=time.perf_counter() - tstart,
                                converged=True, arrays=arr)
    # batches of batch_size trajectories, the reduction after each (1-based
    # indices I), stopping once it reports convergence
    u = prob.u_init if prob.u_init is not None else []
    converged = False
    first_arrays = None
    starts = range(0, N, batch_size)
    for b0 in starts:
        batch = idx[b0:b0 + batch_size]
        outputs = {}
        repeat = 1
        pending = batch
        while len(pending):
            f, u0, p, t0, tf, nvec, probs = _pack(prob, pending, repeat)
            ealg = ensemblealg if (repeat == 1 and len(batch) == N) else \
                _subset_ealg(ensemblealg, pending)
            arr = device_solve(f, *_resident_inputs(ealg, u0, p, t0, tf, nvec), alg, ealg,
                               saveat=sa, abstol=abstol, reltol=reltol, dt=dt, dtmin=dtmin,
                               dtmax=dtmax, maxiters=maxiters, **kw)
            if sa is None:
                # without saveat: (t0, u0) and the final state by the flags
                arr.drop_start, arr.drop_end = save_start is False, save_end is False
            if first_arrays is None:
                first_arrays = arr
            rerun = []
            opts = dict(abstol=abstol, reltol=reltol, dt=dt, dtmin=dtmin, dtmax=dtmax,
                        maxiters=maxiters)
            packed = (f, u0, p, t0, tf, nvec)
            for k, i in enumerate(pending):
                sol = arr.solution(k, probs[k] if probs else None)
                if dense:
                    sol.interp = _dense_interp(packed, k, int(i), alg, ensemblealg, arr.lanes,
                                               opts, kw)
                out, again = prob.output_func(sol, int(i) + 1)
                outputs[int(i)] = out
                if again:
                    rerun.append(i)
            pending = np.array(rerun, dtype=np.int64)
            repeat += 1
        I = range(int(batch[0]) + 1, int(batch[-1]) + 2)
        u, converged = prob.reduction(u, [outputs[int(i)] for i in batch], I)
        converged = bool(converged)
        if converged:
            break
    if len(starts) == 1 and prob.reduction is _default_reduction:
        converged = True  # one batch with the default reduction: complete
    return EnsembleSolution(u=u, elapsedTime=time.perf_counter() - tstart,
                            converged=converged, arrays=first_arrays)
