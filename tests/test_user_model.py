"""Run-time compiled right-hand sides on the CPU: the device compile check
(hipRTC runs without a GPU), and the oracle's user-model path -- the dual-
number Jacobian against analytic and finite-difference Jacobians, and whole
solves of a user model against the same model built in."""
import ctypes as C

import numpy as np
import pytest

import user_model_host as umh

LORENZ = ("du[0] = p[0] * (u[1] - u[0]);\n"
          "du[1] = u[0] * (p[1] - u[2]) - u[1];\n"
          "du[2] = u[0] * u[1] - p[2] * u[2];")
ROBERTSON = ("T y1 = u[0], y2 = u[1], y3 = u[2];\n"
             "T k1 = p[0], k2 = p[1], k3 = p[2];\n"
             "du[0] = -k1 * y1 + k3 * y2 * y3;\n"
             "du[1] = k1 * y1 - k2 * (y2 * y2) - k3 * y2 * y3;\n"
             "du[2] = k2 * (y2 * y2);")
PENDULUM = ("du[0] = u[1];\n"
            "du[1] = -p[0] * sin(u[0]) - p[1] * u[1] + 0.1 * exp(-u[0] * u[0]) * tanh(u[1])"
            " + pow(fabs(u[0]) + 1.0, 1.5) / sqrt(2.0 + u[1] * u[1]) - log(3.0 + cos(u[0]));")


def _mideq():
    import mideq
    return mideq


def test_device_compile_check_and_errors():
    mideq = _mideq()
    assert mideq.UserModel(3, 3, LORENZ).check() == ""
    mideq.UserModel(2, 2, PENDULUM).check()
    with pytest.raises(mideq.MideqError, match="rhs:1:13"):
        mideq.UserModel(2, 1, "du[0] = u[1]\ndu[1] = -u[0];").check()
    with pytest.raises(mideq.MideqError, match="jac"):
        mideq.UserModel(1, 1, "du[0] = -u[0];", jac="J[0][0] = -1.0 +;").check()
    with pytest.raises(ValueError):
        mideq.UserModel(17, 1, "du[0] = 0.0;")


def _jac(lib, n, u, p):
    J = np.zeros(n * n)
    u = np.ascontiguousarray(u, dtype=np.float64)
    p = np.ascontiguousarray(p, dtype=np.float64)
    dp = C.POINTER(C.c_double)
    lib.user_jac(J.ctypes.data_as(dp), u.ctypes.data_as(dp), p.ctypes.data_as(dp), 0.0)
    return J.reshape(n, n)


def _f(lib, u, p):
    du = np.zeros(len(u))
    dp = C.POINTER(C.c_double)
    u = np.ascontiguousarray(u, dtype=np.float64)
    p = np.ascontiguousarray(p, dtype=np.float64)
    lib.user_f64(du.ctypes.data_as(dp), u.ctypes.data_as(dp), p.ctypes.data_as(dp), 0.0)
    return du


def test_dual_jacobian_matches_analytic(oracle):
    """Forward-mode AD of polynomial right-hand sides gives the hand-written
    Jacobians of the built-in models exactly."""
    rng = np.random.default_rng(0)
    for name, rhs in (("lorenz", LORENZ), ("robertson", ROBERTSON)):
        lib = umh.build(3, 3, rhs)
        for _ in range(20):
            u = rng.uniform(-2, 2, 3)
            p = rng.uniform(0.1, 30, 3)
            np.testing.assert_array_equal(_f(lib, u, p), oracle.rhs(name, u, p))
            np.testing.assert_array_equal(_jac(lib, 3, u, p), oracle.jac(name, u, p))


def test_dual_jacobian_transcendental_vs_finite_differences():
    lib = umh.build(2, 2, PENDULUM)
    rng = np.random.default_rng(1)
    for _ in range(20):
        u = rng.uniform(-1.5, 1.5, 2)
        p = rng.uniform(0.5, 5.0, 2)
        J = _jac(lib, 2, u, p)
        Jfd = np.zeros((2, 2))
        for j in range(2):
            h = 1e-6 * max(1.0, abs(u[j]))
            e = np.zeros(2)
            e[j] = h
            Jfd[:, j] = (_f(lib, u + e, p) - _f(lib, u - e, p)) / (2 * h)
        np.testing.assert_allclose(J, Jfd, rtol=1e-7, atol=1e-8)


@pytest.mark.parametrize("alg,precision,controller", [("tsit5", "fp64", 0), ("tsit5", "mixed", 1),
                                                      ("rodas5", "fp64", 0), ("auto", "fp64", 1)])
def test_oracle_user_model_equals_builtin(oracle, alg, precision, controller):
    """The oracle's user-model path (host-compiled statements, dual-number J)
    reproduces the built-in Robertson model bit for bit."""
    umh.install(oracle, 3, 3, ROBERTSON)
    N = 16
    rng = np.random.default_rng(2)
    u0 = np.zeros((3, N))
    u0[0] = 1.0
    p = np.array([0.04, 3e7, 1e4])[:, None] * 10.0 ** rng.uniform(-0.5, 0.5, (3, N))
    kw = dict(alg=alg, precision=precision, controller=controller, abstol=1e-8, reltol=1e-6,
              saveat=np.logspace(-4, 4, 9))
    a = oracle.ensemble("user", u0, p, 0.0, 1e4, **kw)
    b = oracle.ensemble("robertson", u0, p, 0.0, 1e4, **kw)
    for k in ("u", "t", "retcode", "naccept", "nreject", "nf", "saveat"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


BALL = ("du[0] = u[1];\ndu[1] = -p[0];")


def _ball(oracle, alg="tsit5", N=4, tf=6.0):
    from mideq import ContinuousCallback
    cb = ContinuousCallback("u[0]", affect="u[1] = -p[1] * u[1];", direction=-1)
    umh.install(oracle, 2, 2, BALL, callbacks=(cb,))
    h0 = np.array([10.0, 5.0, 2.0, 20.0])[:N]
    u0 = np.vstack([h0, np.zeros(N)])
    p = np.vstack([np.full(N, 9.81), np.full(N, 0.9)])
    return oracle.ensemble("user", u0, p, 0.0, tf, alg=alg, abstol=1e-12, reltol=1e-12), h0


def _bounce_times(h0, g=9.81, e=0.9, tf=6.0):
    t, v, out = np.sqrt(2 * h0 / g), e * np.sqrt(2 * g * h0), []
    while t <= tf:
        out.append(t)
        t += 2 * v / g
        v *= e
    return out


@pytest.mark.parametrize("alg", ["tsit5", "rodas5", "auto"])
def test_oracle_bouncing_ball_event_times(oracle, alg):
    """Downcrossings of the height located on the dense output: the number of
    bounces and the last bounce time follow the closed form."""
    r, h0 = _ball(oracle, alg)
    assert (r["retcode"] == 1).all()
    for i in range(len(h0)):
        times = _bounce_times(h0[i])
        assert r["nevent"][i] == len(times)
        assert abs(r["t_event"][i] - times[-1]) < 1e-8 * max(1.0, times[-1])


def test_oracle_terminate_at_half(oracle):
    """u' = -u from 1, terminate where u = 1/2: t = ln 2, retcode Terminated,
    save points after the event unfilled."""
    from mideq import ContinuousCallback
    umh.install(oracle, 1, 1, "du[0] = -p[0] * u[0];",
                callbacks=(ContinuousCallback("u[0] - 0.5", terminate=True),))
    sa = np.array([0.5, 1.0])
    r = oracle.ensemble("user", np.ones((1, 3)), np.ones((1, 3)), 0.0, 2.0, abstol=1e-12,
                        reltol=1e-12, saveat=sa)
    assert (r["retcode"] == 7).all() and (r["nevent"] == 1).all()
    np.testing.assert_allclose(r["t"], np.log(2.0), rtol=1e-10)
    np.testing.assert_allclose(r["u"][0], 0.5, rtol=1e-10)
    assert np.isnan(r["saveat"][1]).all() and not np.isnan(r["saveat"][0]).any()


def test_oracle_core_jl_callbacks(oracle):
    """test/core.jl:60-79: ContinuousCallback(t - 0.5) and DiscreteCallback(t >= 0.5)
    with no affect on Lorenz both finish with Success; the continuous one
    stops a step at t = 0.5."""
    from mideq import ContinuousCallback, DiscreteCallback
    u0 = np.array([[1.0], [0.0], [0.0]])
    p = np.array([[10.0], [28.0], [8.0 / 3.0]])
    umh.install(oracle, 3, 3, LORENZ, callbacks=(ContinuousCallback("t - 0.5"),))
    r = oracle.ensemble("user", u0, p, 0.0, 1.0, abstol=1e-6, reltol=1e-3)
    assert r["retcode"][0] == 1 and r["nevent"][0] == 1
    assert abs(r["t_event"][0] - 0.5) < 1e-12
    umh.install(oracle, 3, 3, LORENZ, callbacks=(DiscreteCallback("t >= 0.5"),))
    r2 = oracle.ensemble("user", u0, p, 0.0, 1.0, abstol=1e-6, reltol=1e-3)
    assert r2["retcode"][0] == 1 and r2["nevent"][0] >= 1 and r2["t_event"][0] == 1.0
    # without callbacks: the same steps as the built-in model
    umh.install(oracle, 3, 3, LORENZ)
    r3 = oracle.ensemble("user", u0, p, 0.0, 1.0, abstol=1e-6, reltol=1e-3)
    r4 = oracle.ensemble("lorenz", u0, p, 0.0, 1.0, abstol=1e-6, reltol=1e-3)
    np.testing.assert_array_equal(r3["u"], r4["u"])
    assert r3["nevent"][0] == 0


# non-autonomous right-hand sides (ODEProblem f(du, u, p, t)): the time is
# the stage time, Rosenbrock steps add d_i dt df/dt with df/dt by dual numbers
PROTHERO = "du[0] = p[0] * (u[0] - sin(t)) + cos(t);"
FORCED = "du[0] = u[1];\ndu[1] = -p[0] * p[0] * u[0] + t;"


def test_time_dependence_detected():
    mideq = _mideq()
    assert mideq.UserModel(1, 1, PROTHERO).check() == ""  # f(u, p, t), jac and tgrad compile
    assert mideq.UserModel(2, 1, FORCED, jac="J[0][0] = 0.0; J[0][1] = 1.0;\n"
                           "J[1][0] = -p[0] * p[0] + 0.0 * t; J[1][1] = 0.0;").check() == ""
    assert umh.uses_time(PROTHERO) and umh.uses_time(FORCED)
    assert not umh.uses_time(LORENZ) and not umh.uses_time(ROBERTSON)
    assert not umh.uses_time("T tt = u[0]; du[0] = tt * sqrt(u[0]);")
    assert umh.uses_time("du[0] = (t);")


@pytest.mark.parametrize("alg", ["rodas5", "rodas5p", "tsit5"])
def test_oracle_user_prothero_equals_builtin(oracle, alg):
    """The user form of y' = lambda (y - sin t) + cos t equals the oracle's
    built-in one bit for bit: same RHS expression, J = lambda exactly from
    the dual numbers, and df/dt from Dual<1> rounding like the analytic
    -lambda cos t - sin t; its solution is sin t (y(0) = 0)."""
    umh.install(oracle, 1, 1, PROTHERO)
    N = 64
    lam = -np.logspace(1, 5, N)
    kw = dict(alg=alg, abstol=1e-9, reltol=1e-7)
    a = oracle.ensemble("user", np.zeros((1, N)), lam[None, :], 0.0, 3.0, **kw)
    b = oracle.ensemble("prothero", np.zeros((1, N)), lam[None, :], 0.0, 3.0, **kw)
    assert (a["retcode"] == 1).all()
    for k in ("u", "t", "naccept", "nreject", "nf"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    np.testing.assert_allclose(a["u"][0], np.sin(3.0), atol=1e-6)


@pytest.mark.parametrize("alg", ["tsit5", "rodas5"])
def test_oracle_user_forced_oscillator_closed_form(oracle, alg):
    """x'' = -w^2 x + t from rest: x = t / w^2 - sin(w t) / w^3."""
    umh.install(oracle, 2, 1, FORCED)
    N = 32
    w = np.linspace(0.5, 3.0, N)
    r = oracle.ensemble("user", np.zeros((2, N)), w[None, :], 0.0, 4.0, alg=alg,
                        abstol=1e-11, reltol=1e-11)
    assert (r["retcode"] == 1).all()
    x = 4.0 / w**2 - np.sin(4.0 * w) / w**3
    np.testing.assert_allclose(r["u"][0], x, rtol=1e-7, atol=1e-9)


def test_uses_time_rule_and_local_t():
    """The library's rule for a non-autonomous model (mideq_uses_time) skips
    comments and literals and equals the host restatement; a model declaring
    its own local t compiles (the statements are a block of their own)."""
    mideq = _mideq()
    from mideq import _abi
    lib = _abi.load()
    lib.mideq_uses_time.argtypes = [C.c_char_p]
    cases = {
        PROTHERO: 1, FORCED: 1, LORENZ: 0, ROBERTSON: 0,
        "du[0] = -u[0]; // decays in t": 0,
        "du[0] = -u[0]; /* not t-dependent:\n t */": 0,
        "du[0] = -u[0] * (t > 1.0 ? 2.0 : 1.0);": 1,
        "const char* s = \"t\"; du[0] = -u[0]; char c = 't';": 0,
        "T tt = u[0]; du[0] = tt;": 0,
        "du[0] = s.t;": 0,
        "du[0] = sin(t) /* t */;": 1,
        "double t = u[0]; du[0] = -t;": 1,
    }
    for text, want in cases.items():
        assert lib.mideq_uses_time(text.encode()) == want, text
        assert umh.uses_time(text) == bool(want), text
    assert mideq.UserModel(1, 1, "T t = u[0]; du[0] = -t;").check() == ""


@pytest.mark.parametrize("saves,per_event", [((True, True), 2), ((False, True), 1),
                                             ((True, False), 1)])
def test_oracle_event_points_bouncing_ball(oracle, saves, per_event):
    """save_positions at the bounces of a ball: per bounce the state just
    before the affect (height ~0, falling at v) and / or just after it
    (the same height, rising at e v), at the closed-form bounce times; a
    capacity smaller than the count keeps the first points and reports
    them all."""
    from mideq import ContinuousCallback
    e, g = 0.9, 9.81
    cb = ContinuousCallback("u[0]", affect="u[1] = -p[1] * u[1];", direction=-1,
                            save_positions=saves)
    umh.install(oracle, 2, 2, BALL, callbacks=(cb,))
    h0 = np.array([10.0, 5.0, 2.0, 20.0])
    u0 = np.vstack([h0, np.zeros(4)])
    p = np.vstack([np.full(4, g), np.full(4, e)])
    r = oracle.ensemble("user", u0, p, 0.0, 6.0, abstol=1e-12, reltol=1e-12, event_cap=64)
    np.testing.assert_array_equal(r["ev_n"], per_event * r["nevent"])
    for i in range(4):
        times = _bounce_times(h0[i], g=g, e=e)
        m = int(r["ev_n"][i])
        et, eu = r["ev_t"][i, :m], r["ev_u"][i, :m]
        assert np.isnan(r["ev_t"][i, m:]).all()
        np.testing.assert_allclose(et[::per_event], times, rtol=1e-8)
        assert np.abs(eu[:, 0]).max() < 1e-6  # on the floor
        if saves == (True, True):
            np.testing.assert_array_equal(et[0::2], et[1::2])
            np.testing.assert_array_equal(eu[1::2, 1], -e * eu[0::2, 1])
            assert (eu[0::2, 1] < 0).all() and (eu[1::2, 1] > 0).all()
        elif saves == (False, True):
            assert (eu[:, 1] > 0).all()
        else:
            assert (eu[:, 1] < 0).all()
    small = oracle.ensemble("user", u0, p, 0.0, 6.0, abstol=1e-12, reltol=1e-12, event_cap=2)
    np.testing.assert_array_equal(small["ev_n"], r["ev_n"])
    np.testing.assert_array_equal(small["ev_t"], r["ev_t"][:, :2])
    np.testing.assert_array_equal(small["ev_u"], r["ev_u"][:, :2])


def test_oracle_save_everystep_closed_form(oracle):
    """save_everystep on u' = -k u: one point per accepted step, strictly
    increasing times ending at tf, states on exp(-k t) to the tolerance; with
    saveat holding tf the last step's end is not saved a second time, and a
    callback step saves its end once (the point before the affect is that
    one) plus the point after the affect."""
    from mideq import DiscreteCallback
    rhs = "du[0] = -p[0] * u[0];"
    umh.install(oracle, 1, 1, rhs, save_everystep=True)
    k = np.array([0.5, 1.0, 2.0, 4.0])
    u0, p = np.ones((1, 4)), k[None, :]
    r = oracle.ensemble("user", u0, p, 0.0, 2.0, abstol=1e-11, reltol=1e-11, event_cap=256)
    assert (r["retcode"] == 1).all()
    np.testing.assert_array_equal(r["ev_n"], r["naccept"])
    for i in range(4):
        m = int(r["ev_n"][i])
        t = r["ev_t"][i, :m]
        assert (np.diff(t) > 0).all() and t[-1] == 2.0 and np.isnan(r["ev_t"][i, m:]).all()
        np.testing.assert_allclose(r["ev_u"][i, :m, 0], np.exp(-k[i] * t), rtol=1e-9)
        assert r["ev_u"][i, m - 1, 0] == r["u"][0, i]
    r2 = oracle.ensemble("user", u0, p, 0.0, 2.0, abstol=1e-11, reltol=1e-11, event_cap=256,
                         saveat=np.array([1.0, 2.0]))
    np.testing.assert_array_equal(r2["ev_n"], r2["naccept"] - 1)
    np.testing.assert_array_equal(r2["ev_t"][:, :int(r2["ev_n"][0])],
                                  r["ev_t"][:, :int(r2["ev_n"][0])])
    # a discrete callback halving u at every step end past t = 1: each such
    # step keeps its end before the affect and the state after it
    cb = DiscreteCallback("t > 1.0", affect="u[0] = 0.5 * u[0];")
    umh.install(oracle, 1, 1, rhs, callbacks=(cb,), save_everystep=True)
    r3 = oracle.ensemble("user", u0, p, 0.0, 2.0, abstol=1e-11, reltol=1e-11, event_cap=512)
    np.testing.assert_array_equal(r3["ev_n"], r3["naccept"] + r3["nevent"])
    for i in range(4):
        m = int(r3["ev_n"][i])
        t, u = r3["ev_t"][i, :m], r3["ev_u"][i, :m, 0]
        late = np.nonzero(t > 1.0)[0]
        # past t = 1 the points come in pairs (before, after) at one time
        np.testing.assert_array_equal(t[late[0::2]], t[late[1::2]])
        np.testing.assert_array_equal(u[late[1::2]], 0.5 * u[late[0::2]])


def test_oracle_continuous_callback_without_affect_saves_before_only(oracle):
    """A continuous callback with no affect (upstream's affect! = nothing)
    leaves u unmodified, so of save_positions (true, true) only the point
    before it is saved; a discrete one saves both (its affect always runs)."""
    from mideq import ContinuousCallback, DiscreteCallback
    rhs = "du[0] = -p[0] * u[0];"
    u0, p = np.ones((1, 3)), np.array([[0.5, 1.0, 2.0]])
    cb = ContinuousCallback("t - 0.5")
    umh.install(oracle, 1, 1, rhs, callbacks=(cb,))
    r = oracle.ensemble("user", u0, p, 0.0, 1.0, abstol=1e-10, reltol=1e-10, event_cap=8)
    np.testing.assert_array_equal(r["ev_n"], [1, 1, 1])
    np.testing.assert_allclose(r["ev_t"][:, 0], 0.5, atol=1e-12)
    np.testing.assert_allclose(r["ev_u"][:, 0, 0], np.exp(-0.5 * p[0]), rtol=1e-8)
    dcb = DiscreteCallback("t >= 0.9")
    umh.install(oracle, 1, 1, rhs, callbacks=(dcb,))
    r2 = oracle.ensemble("user", u0, p, 0.0, 1.0, abstol=1e-10, reltol=1e-10, event_cap=64)
    np.testing.assert_array_equal(r2["ev_n"], 2 * r2["nevent"])


def test_oracle_tstops_land_and_dose(oracle):
    """tstops: the steps land exactly on the ones inside the span (kept as
    step ends with save_everystep; ones outside it are ignored), and a
    discrete callback at t == 0.5 with tstops [0.5] -- a dose -- gives the
    closed form u(1) = (exp(-k/2) + 1) exp(-k/2) of u' = -k u."""
    from mideq import DiscreteCallback
    rhs = "du[0] = -p[0] * u[0];"
    k = np.array([0.5, 1.0, 3.0])
    u0, p = np.ones((1, 3)), k[None, :]
    umh.install(oracle, 1, 1, rhs, save_everystep=True)
    ts = np.array([-1.0, 0.3, 0.77, 5.0])
    r = oracle.ensemble("user", u0, p, 0.0, 1.0, abstol=1e-10, reltol=1e-10, event_cap=512,
                        tstops=ts)
    for i in range(3):
        t = r["ev_t"][i, :int(r["ev_n"][i])]
        assert 0.3 in t and 0.77 in t and t[-1] == 1.0
    np.testing.assert_allclose(r["u"][0], np.exp(-k), rtol=1e-8)
    dose = DiscreteCallback("t == 0.5", affect="u[0] = u[0] + 1.0;")
    umh.install(oracle, 1, 1, rhs, callbacks=(dose,))
    r2 = oracle.ensemble("user", u0, p, 0.0, 1.0, abstol=1e-10, reltol=1e-10,
                         tstops=np.array([0.5]))
    np.testing.assert_array_equal(r2["nevent"], [1, 1, 1])
    np.testing.assert_allclose(r2["u"][0], (np.exp(-k / 2) + 1.0) * np.exp(-k / 2), rtol=1e-8)
    r3 = oracle.ensemble("user", u0, p, 0.0, 1.0, abstol=1e-10, reltol=1e-10)
    assert (r3["nevent"] == 0).all()  # without the tstop no step ends at 0.5
