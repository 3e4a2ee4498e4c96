"""ctypes binding of the CPU oracle (oracle/_build/libmideq_oracle.so).

Test infrastructure: the oracle is the checker, never the thing measured.
"""
import ctypes as C
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# MIDEQ_ORACLE_LIB selects another build of the same source (the sanitizer build)
LIB = os.environ.get("MIDEQ_ORACLE_LIB") or os.path.join(REPO, "oracle", "_build",
                                                         "libmideq_oracle.so")

MODEL = dict(exp=0, lorenz=1, robertson=2, vdp=3, brusselator=4, linear=5, prothero=6, user=7)
ALG = dict(tsit5=0, rodas5=1, auto=2, rodas5p=3, auto5p=4)
PREC = dict(fp64=0, mixed=1)
RC = dict(default=0, success=1, maxiters=2, dtlessthanmin=3, unstable=4, dtnan=5,
          initialfailure=6, terminated=7)
NMAX = 128

_dp = C.POINTER(C.c_double)
TRACE_DTYPE = np.dtype([("t", "<f8"), ("dt_hdr", "<f8"), ("dt", "<f8"), ("eig", "<f8"),
                        ("mode", "<i4"), ("accepted", "<i4")])
_ip = C.POINTER(C.c_int32)


class Options(C.Structure):
    _fields_ = [
        ("alg", C.c_int32),
        ("precision", C.c_int32),
        ("abstol", C.c_double),
        ("reltol", C.c_double),
        ("dt", C.c_double),
        ("dtmin", C.c_double),
        ("dtmax", C.c_double),
        ("maxiters", C.c_int64),
        ("saveat", _dp),
        ("ns", C.c_int32),
        ("qsteady_default", C.c_int32),
        ("controller", C.c_int32),
        ("norm_lanes", C.c_int32),
        ("dt0_in", _dp),
        ("alg_hint", C.POINTER(C.c_int8)),
        ("save_layout", C.c_int32),
        ("ev_t", _dp),
        ("ev_u", _dp),
        ("ev_n", C.POINTER(C.c_int32)),
        ("ev_cap", C.c_int32),
        ("reserved", C.c_int32),
        ("tstops", _dp),
        ("ntstops", C.c_int32),
        ("reserved2", C.c_int32),
    ]


class Result(C.Structure):
    _fields_ = [
        ("u", C.c_double * NMAX),
        ("t", C.c_double),
        ("retcode", C.c_int32),
        ("naccept", C.c_int32),
        ("nreject", C.c_int32),
        ("nf", C.c_int32),
        ("njac", C.c_int32),
        ("dt0", C.c_double),
        ("dt_first_accept", C.c_double),
        ("nstiff_steps", C.c_int32),
        ("nswitch", C.c_int32),
        ("first_stiff_iter", C.c_int32),
        ("nevent", C.c_int32),
        ("t_event", C.c_double),
        ("nevsave", C.c_int32),
        ("reserved", C.c_int32),
    ]


def _ptr(a, ct=C.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ct))


class Oracle:
    def __init__(self, path=LIB):
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run make -C oracle)")
        self.lib = L = C.CDLL(path)
        L.oracle_solve.argtypes = [C.c_int32, C.c_int32, _dp, _dp, C.c_int32, C.c_double,
                                   C.c_double, C.POINTER(Options), _dp, C.POINTER(Result)]
        L.oracle_ensemble_solve.argtypes = [
            C.c_int32, C.c_int32, C.c_int32, C.c_int64, _dp, _dp, _dp, _dp, _ip,
            C.POINTER(Options), C.c_int32, _dp, _dp, _dp, C.POINTER(C.c_int8), _ip, _ip, _ip,
            _dp, _ip, _dp, _ip, _dp]
        L.oracle_tsit5_tableau.argtypes = [_dp, _dp, _dp, _dp]
        L.oracle_rodas5_tableau.argtypes = [_dp, _dp, _dp, _dp, _dp]
        L.oracle_rosenbrock_tableau.argtypes = [C.c_int32, _dp, _dp, _dp, _dp, _dp]
        L.oracle_rosenbrock_dense.argtypes = [C.c_int32, _dp]
        L.oracle_rosenbrock_interp.argtypes = [C.c_int32, C.c_int32, C.c_int32, _dp, _dp,
                                               C.c_double, C.c_double, C.c_int32, _dp, _dp, _dp]
        for nm in ("oracle_ctl_log2", "oracle_ctl_exp2"):
            getattr(L, nm).argtypes = [C.c_double]
            getattr(L, nm).restype = C.c_double
        L.oracle_fast_log2.argtypes = [C.c_double]
        L.oracle_fast_log2.restype = C.c_double
        L.oracle_fast_exp2.argtypes = [C.c_double]
        L.oracle_fast_exp2.restype = C.c_double
        L.oracle_ctl_pow.argtypes = [C.c_double, C.c_double]
        L.oracle_ctl_pow.restype = C.c_double
        L.oracle_controller.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, _dp, _dp]
        L.oracle_lu_solve.argtypes = [C.c_int32, _dp, _dp]
        L.oracle_rhs.argtypes = [C.c_int32, C.c_int32, _dp, _dp, C.c_double, _dp]
        L.oracle_jac.argtypes = [C.c_int32, C.c_int32, _dp, _dp, C.c_double, _dp]
        L.oracle_initdt.argtypes = [C.c_int32, C.c_int32, _dp, _dp, C.c_double, C.c_double,
                                    C.POINTER(Options), C.c_int]
        L.oracle_initdt.restype = C.c_double
        L.oracle_tsit5_step.argtypes = [C.c_int32, C.c_int32, _dp, _dp, C.c_double, C.c_double,
                                        C.c_double, C.c_double, _dp, _dp, _dp]
        L.oracle_rodas5_step.argtypes = [C.c_int32, C.c_int32, _dp, _dp, C.c_double,
                                         C.c_double, C.c_double, C.c_double, _dp, _dp]
        L.oracle_rosenbrock_step.argtypes = [C.c_int32] + L.oracle_rodas5_step.argtypes
        L.oracle_tsit5_interp.argtypes = [C.c_int32, _dp, _dp, C.c_double, C.c_double, _dp]

    # ---------------------------------------------------------------- options
    @staticmethod
    def options(alg="tsit5", precision="fp64", abstol=1e-6, reltol=1e-3, dt=0.0, dtmin=0.0,
                dtmax=0.0, maxiters=0, saveat=None, qsteady_default=1, norm_lanes=64,
                controller=0, dt0=None, alg_hint=None, save_layout=0, tstops=None):
        o = Options()
        o.save_layout = save_layout
        if tstops is not None:
            o._tstops_keep = np.sort(np.ascontiguousarray(tstops, dtype=np.float64))
            o.tstops = _ptr(o._tstops_keep)
            o.ntstops = len(o._tstops_keep)
        o.alg = ALG[alg]
        o.precision = PREC[precision]
        o.abstol = abstol
        o.reltol = reltol
        o.dt = dt
        o.dtmin = dtmin
        o.dtmax = dtmax
        o.maxiters = maxiters
        o.qsteady_default = qsteady_default
        o.norm_lanes = norm_lanes
        o.controller = controller
        if dt0 is not None:
            o._dt0_keep = np.ascontiguousarray(dt0, dtype=np.float64)
            o.dt0_in = _ptr(o._dt0_keep)
        if alg_hint is not None:
            o._hint_keep = np.ascontiguousarray(alg_hint, dtype=np.int8)
            o.alg_hint = o._hint_keep.ctypes.data_as(C.POINTER(C.c_int8))
        if saveat is not None:
            sa = np.ascontiguousarray(saveat, dtype=np.float64)
            o._saveat_keep = sa
            o.saveat = _ptr(sa)
            o.ns = len(sa)
        else:
            o.saveat = None
            o.ns = 0
        return o

    # ------------------------------------------------------------- one solve
    def solve(self, model, u0, p, tspan, **kw):
        u0 = np.ascontiguousarray(u0, dtype=np.float64).reshape(-1)
        p = np.ascontiguousarray(p, dtype=np.float64).reshape(-1)
        o = self.options(**kw)
        n = len(u0)
        save = None
        if o.ns:
            save = np.zeros((o.ns, n))
        r = Result()
        rc = self.lib.oracle_solve(MODEL[model], n, _ptr(u0), _ptr(p), len(p), tspan[0],
                                   tspan[1], C.byref(o), _ptr(save), C.byref(r))
        assert rc == 0
        out = dict(u=np.array(r.u[:n]), t=r.t, retcode=r.retcode, naccept=r.naccept,
                   nreject=r.nreject, nf=r.nf, njac=r.njac, dt0=r.dt0,
                   dt_first_accept=r.dt_first_accept, nstiff=r.nstiff_steps,
                   nswitch=r.nswitch, first_stiff_iter=r.first_stiff_iter)
        if save is not None:
            out["saveat"] = save
        return out

    def solve_trace(self, model, u0, p, tspan, cap=200000, **kw):
        """oracle_solve with the record of every attempted step: (result
        dict, structured array of t, dt_hdr, dt, eig, mode, accepted)."""
        u0 = np.ascontiguousarray(u0, dtype=np.float64).reshape(-1)
        p = np.ascontiguousarray(p, dtype=np.float64).reshape(-1)
        o = self.options(**kw)
        n = len(u0)
        rec = np.zeros(cap, dtype=TRACE_DTYPE)
        r = Result()
        L = self.lib
        L.oracle_solve_trace.argtypes = [C.c_int32, C.c_int32, _dp, _dp, C.c_int32, C.c_double,
                                         C.c_double, C.POINTER(Options), _dp, C.POINTER(Result),
                                         C.c_void_p, C.c_int32]
        m = L.oracle_solve_trace(MODEL[model], n, _ptr(u0), _ptr(p), len(p), tspan[0], tspan[1],
                                 C.byref(o), None, C.byref(r), rec.ctypes.data, cap)
        assert 0 <= m <= cap, m
        out = dict(u=np.array(r.u[:n]), t=r.t, retcode=r.retcode, naccept=r.naccept,
                   nreject=r.nreject, nf=r.nf, nstiff=r.nstiff_steps, nswitch=r.nswitch,
                   first_stiff_iter=r.first_stiff_iter, dt0=r.dt0)
        return out, rec[:m]

    # ------------------------------------------------------ ensemble (SoA)
    def ensemble(self, model, u0, p, t0, tf, nthreads=1, nvec=None, event_cap=0, **kw):
        """u0: [n, N], p: [np, N], t0/tf: [N] or scalars.  Returns dict of SoA arrays.
        event_cap > 0: the points saved at callback events too (ev_t [N, cap],
        ev_u [N, cap, n], ev_n [N], NaN past the stored ones)."""
        u0 = np.ascontiguousarray(u0, dtype=np.float64)
        p = np.ascontiguousarray(p, dtype=np.float64)
        n, N = u0.shape
        t0 = np.ascontiguousarray(np.broadcast_to(t0, (N,)), dtype=np.float64)
        tf = np.ascontiguousarray(np.broadcast_to(tf, (N,)), dtype=np.float64)
        o = self.options(**kw)
        # saves: [ns][n][N] (save_layout 0) or [N][ns][n] (1); rows of a
        # ragged instance past its size stay NaN
        if not o.ns:
            save = None
        elif o.save_layout == 1:
            save = np.full((N, o.ns, n), np.nan)
        else:
            save = np.full((o.ns, n, N), np.nan)
        uf = np.zeros((n, N))
        tfin = np.zeros(N)
        rc = np.zeros(N, dtype=np.int8)
        na = np.zeros(N, dtype=np.int32)
        nr = np.zeros(N, dtype=np.int32)
        nf = np.zeros(N, dtype=np.int32)
        dt0 = np.zeros(N)
        nst = np.zeros(N, dtype=np.int32)
        dtf = np.zeros(N)
        nev = np.zeros(N, dtype=np.int32)
        tev = np.zeros(N)
        nv = None if nvec is None else np.ascontiguousarray(nvec, dtype=np.int32)
        if event_cap > 0:
            ev_t = np.full((N, event_cap), np.nan)
            ev_u = np.full((N, event_cap, n), np.nan)
            ev_n = np.zeros(N, dtype=np.int32)
            o.ev_t, o.ev_u = _ptr(ev_t), _ptr(ev_u)
            o.ev_n, o.ev_cap = _ptr(ev_n, C.c_int32), event_cap
        r = self.lib.oracle_ensemble_solve(
            MODEL[model], n, p.shape[0], N, _ptr(u0), _ptr(p), _ptr(t0), _ptr(tf),
            _ptr(nv, C.c_int32), C.byref(o), nthreads, _ptr(save), _ptr(uf), _ptr(tfin),
            rc.ctypes.data_as(C.POINTER(C.c_int8)), _ptr(na, C.c_int32), _ptr(nr, C.c_int32),
            _ptr(nf, C.c_int32), _ptr(dt0), _ptr(nst, C.c_int32), _ptr(dtf),
            _ptr(nev, C.c_int32), _ptr(tev))
        assert r == 0
        out = dict(u=uf, t=tfin, retcode=rc, naccept=na, nreject=nr, nf=nf, dt0=dt0,
                   nstiff=nst, dt_first=dtf, nevent=nev, t_event=tev)
        if save is not None:
            out["saveat"] = save
        if event_cap > 0:
            out.update(ev_t=ev_t, ev_u=ev_u, ev_n=ev_n)
        return out

    # ------------------------------------------------------------ internals
    def tsit5_tableau(self):
        c = np.zeros(7)
        a = np.zeros(49)
        bt = np.zeros(7)
        r = np.zeros(28)
        self.lib.oracle_tsit5_tableau(_ptr(c), _ptr(a), _ptr(bt), _ptr(r))
        return c, a.reshape(7, 7), bt, r.reshape(7, 4)

    def rodas5_tableau(self, which=0):
        """which: 0 Rodas5, 1 Rodas5P."""
        g = C.c_double()
        a = np.zeros(64)
        Cm = np.zeros(64)
        c = np.zeros(8)
        d = np.zeros(8)
        self.lib.oracle_rosenbrock_tableau(which, C.byref(g), _ptr(a), _ptr(Cm), _ptr(c), _ptr(d))
        return g.value, a.reshape(8, 8), Cm.reshape(8, 8), c, d

    def rosenbrock_dense(self, which=0):
        """Dense-output coefficients H [3, 8] (the error stage column is 0)."""
        h = np.zeros(21)
        self.lib.oracle_rosenbrock_dense(which, _ptr(h))
        return np.concatenate([h.reshape(3, 7), np.zeros((3, 1))], axis=1)

    def rosenbrock_interp(self, which, model, u, p, t, dt, thetas):
        """One FP64 Rosenbrock step, then its dense output at each theta:
        (u1, values [len(thetas), n])."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        p = np.ascontiguousarray(p, dtype=np.float64)
        th = np.ascontiguousarray(thetas, dtype=np.float64)
        u1 = np.zeros_like(u)
        out = np.zeros((len(th), len(u)))
        rc = self.lib.oracle_rosenbrock_interp(which, MODEL[model], len(u), _ptr(u), _ptr(p), t,
                                               dt, len(th), _ptr(th), _ptr(u1), _ptr(out))
        assert rc == 0
        return u1, out

    def lu_solve(self, A, b):
        A = np.ascontiguousarray(A, dtype=np.float64).copy()
        b = np.ascontiguousarray(b, dtype=np.float64).copy()
        rc = self.lib.oracle_lu_solve(len(b), _ptr(A), _ptr(b))
        return rc, b

    def rhs(self, model, u, p, t=0.0):
        u = np.ascontiguousarray(u, dtype=np.float64)
        p = np.ascontiguousarray(p, dtype=np.float64)
        du = np.zeros_like(u)
        self.lib.oracle_rhs(MODEL[model], len(u), _ptr(u), _ptr(p), t, _ptr(du))
        return du

    def jac(self, model, u, p, t=0.0):
        u = np.ascontiguousarray(u, dtype=np.float64)
        p = np.ascontiguousarray(p, dtype=np.float64)
        n = len(u)
        J = np.zeros(n * n)
        self.lib.oracle_jac(MODEL[model], n, _ptr(u), _ptr(p), t, _ptr(J))
        return J.reshape(n, n)

    def initdt(self, model, u0, p, tspan, order=5, **kw):
        u0 = np.ascontiguousarray(u0, dtype=np.float64)
        p = np.ascontiguousarray(p, dtype=np.float64)
        o = self.options(**kw)
        return self.lib.oracle_initdt(MODEL[model], len(u0), _ptr(u0), _ptr(p), tspan[0],
                                      tspan[1], C.byref(o), order)

    def tsit5_step(self, model, u, p, t, dt, abstol, reltol):
        u = np.ascontiguousarray(u, dtype=np.float64)
        p = np.ascontiguousarray(p, dtype=np.float64)
        n = len(u)
        u1 = np.zeros(n)
        e = C.c_double()
        k = np.zeros(7 * n)
        self.lib.oracle_tsit5_step(MODEL[model], n, _ptr(u), _ptr(p), t, dt, abstol, reltol,
                                   _ptr(u1), C.byref(e), _ptr(k))
        return u1, e.value, k.reshape(7, n)

    def rodas5_step(self, model, u, p, t, dt, abstol, reltol, which=0):
        """One fixed Rosenbrock step; which: 0 Rodas5, 1 Rodas5P."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        p = np.ascontiguousarray(p, dtype=np.float64)
        n = len(u)
        u1 = np.zeros(n)
        e = C.c_double()
        rc = self.lib.oracle_rosenbrock_step(which, MODEL[model], n, _ptr(u), _ptr(p), t, dt,
                                             abstol, reltol, _ptr(u1), C.byref(e))
        return rc, u1, e.value

    def tsit5_interp(self, uprev, k, dt, theta):
        uprev = np.ascontiguousarray(uprev, dtype=np.float64)
        k = np.ascontiguousarray(k, dtype=np.float64)
        out = np.zeros_like(uprev)
        self.lib.oracle_tsit5_interp(len(uprev), _ptr(uprev), _ptr(k), dt, theta, _ptr(out))
        return out

    def controller(self, EEst, qold, dt, qsteady_max=1.0):
        dtn = C.c_double()
        qn = C.c_double()
        acc = self.lib.oracle_controller(EEst, qold, dt, qsteady_max, C.byref(dtn), C.byref(qn))
        return bool(acc), dtn.value, qn.value

    def ctl_pow(self, x, y):
        return self.lib.oracle_ctl_pow(x, y)

    def ctl_log2(self, x):
        return self.lib.oracle_ctl_log2(x)

    def ctl_exp2(self, x):
        return self.lib.oracle_ctl_exp2(x)

    def fast_log2(self, x):
        return self.lib.oracle_fast_log2(x)

    def fast_exp2(self, x):
        return self.lib.oracle_fast_exp2(x)


_INST = None


def load():
    global _INST
    if _INST is None:
        _INST = Oracle()
    return _INST
