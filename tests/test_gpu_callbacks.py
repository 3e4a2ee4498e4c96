"""In-kernel callbacks (ContinuousCallback / DiscreteCallback, qa.jl:26,36;
test/core.jl:60-79) compiled into the right-hand side: event location on
the dense output (Tsit5 and Rosenbrock), affects, terminate!, saveat around
events -- bit-exact against the oracle running the same statements, and
against closed forms."""
import numpy as np
import pytest

import user_model_host as umh
from gpu_util import ALGS, assert_identical, cpu, say
from mideq import (CallbackSet, ContinuousCallback, DiscreteCallback, EnsembleMI355X,
                   UserModel, device_solve, lorenz, robertson, workloads)
from mideq.scimlbase import BUILTIN_SOURCE
from test_user_model import BALL, LORENZ, _bounce_times

pytestmark = pytest.mark.gpu


def run(f, u0, p, t0, tf, alg="tsit5", precision="fp64", schedule="lane", controller="reference",
        saveat=None, callback=None, **kw):
    N = u0.shape[1]
    t0 = np.broadcast_to(np.asarray(t0, dtype=np.float64), (N,)).copy()
    tf = np.broadcast_to(np.asarray(tf, dtype=np.float64), (N,)).copy()
    ealg = EnsembleMI355X(precision=precision, schedule=schedule, controller=controller)
    return device_solve(f, np.ascontiguousarray(u0), np.ascontiguousarray(p), t0, tf, None,
                        ALGS[alg], ealg, saveat=saveat, callback=callback, **kw)


def _check(g, c):
    assert_identical(g, c)
    np.testing.assert_array_equal(g.nevent, c["nevent"])
    np.testing.assert_array_equal(g.t_event, c["t_event"])


def test_core_jl_callbacks_on_lorenz(oracle):
    """test/core.jl:60-79 on a 2,000-trajectory Lorenz ensemble (the built-in
    model with the callbacks compiled in).  The continuous callback has no
    affect, so of its default save_positions (true, true) only the point
    before it is saved (u is unmodified: upstream saves nothing after)."""
    N = 2000
    say(f"core.jl callbacks lorenz N={N}")
    w = workloads.lorenz(N, seed=111)
    kw = dict(abstol=1e-6, reltol=1e-3)
    cb = ContinuousCallback("t - 0.5")
    g = run(lorenz, w["u0"], w["p"], 0.0, 1.0, callback=cb, **kw)
    assert (g.retcode == 1).all() and (g.nevent == 1).all()
    assert np.abs(g.t_event - 0.5).max() < 1e-12
    umh.install(oracle, 3, 3, BUILTIN_SOURCE["lorenz"][2], callbacks=(cb,))
    c = cpu(oracle, "user", w["u0"], w["p"], 0.0, 1.0, event_cap=16, **kw)
    _check(g, c)
    assert (g.ev_n == 1).all() and (g.ev_t[:, 0] == g.t_event).all()
    np.testing.assert_array_equal(g.ev_n, c["ev_n"])
    np.testing.assert_array_equal(g.ev_u, c["ev_u"])
    dcb = DiscreteCallback("t >= 0.5")
    g2 = run(lorenz, w["u0"], w["p"], 0.0, 1.0, callback=dcb, **kw)
    assert (g2.retcode == 1).all() and (g2.nevent >= 1).all() and (g2.t_event == 1.0).all()
    umh.install(oracle, 3, 3, BUILTIN_SOURCE["lorenz"][2], callbacks=(dcb,))
    _check(g2, cpu(oracle, "user", w["u0"], w["p"], 0.0, 1.0, **kw))


@pytest.mark.parametrize("alg,precision,schedule", [("tsit5", "fp64", "lane"),
                                                    ("tsit5", "mixed", "persistent"),
                                                    ("rodas5", "fp64", "lane"),
                                                    ("auto5p", "fp64", "lane")])
def test_bouncing_ball(oracle, alg, precision, schedule):
    """Downcrossings of the height with a restitution affect: bit-exact vs
    the oracle, bounce count and times vs the closed form, saveat values on
    both sides of the bounces."""
    N = 1024
    say(f"bouncing ball {alg} {precision} {schedule} N={N}")
    cb = ContinuousCallback("u[0]", affect="u[1] = -p[1] * u[1];", direction=-1)
    model = UserModel(2, 2, BALL, callbacks=(cb,))
    rng = np.random.default_rng(112)
    # no Zeno point before tf and no bounce shorter than a tenth of the
    # longest step (dtmax = tf)
    h0 = rng.uniform(5.0, 20.0, N)
    u0 = np.vstack([h0, np.zeros(N)])
    p = np.vstack([np.full(N, 9.81), rng.uniform(0.8, 0.95, N)])
    sa = np.linspace(0.0, 6.0, 25)
    kw = dict(abstol=1e-10, reltol=1e-10, saveat=sa)
    g = run(model, u0, p, 0.0, 6.0, alg=alg, precision=precision, schedule=schedule, **kw)
    umh.install(oracle, 2, 2, BALL, callbacks=(cb,))
    c = cpu(oracle, "user", u0, p, 0.0, 6.0, alg=alg, precision=precision, **kw)
    assert (c["retcode"] == 1).all()
    _check(g, c)
    if precision == "fp64":
        for i in range(0, N, 97):
            times = _bounce_times(h0[i], e=p[1, i])
            assert g.nevent[i] == len(times)
            assert abs(g.t_event[i] - times[-1]) < 1e-7 * max(1.0, times[-1])
        assert (g.us[:, 0, :] > -1e-6).all()  # the ball stays above the floor


@pytest.mark.parametrize("interp_points", [10, 0])
def test_terminate_and_callback_set(oracle, interp_points):
    """terminate! at u = 1/2 of u' = -k u (t = ln 2 / k) with a discrete
    callback in the same set; saves after termination stay NaN.  With
    interp_points = 0 only step ends are checked (a monotone condition)."""
    N = 512
    say(f"terminate + CallbackSet interp_points={interp_points} N={N}")
    cbs = CallbackSet(ContinuousCallback("u[0] - 0.5", terminate=True,
                                         interp_points=interp_points),
                      DiscreteCallback("t > 0.25 && t < 0.3", affect="u[0] = u[0] * 1.0;"))
    model = UserModel(1, 1, "du[0] = -p[0] * u[0];", callbacks=cbs)
    rng = np.random.default_rng(113)
    k = rng.uniform(0.5, 3.0, N)
    sa = np.linspace(0.0, 2.0, 9)
    kw = dict(abstol=1e-12, reltol=1e-12, saveat=sa)
    g = run(model, np.ones((1, N)), k[None, :], 0.0, 4.0, **kw)
    assert (g.retcode == 7).all()
    np.testing.assert_allclose(g.t_final, np.log(2.0) / k, rtol=1e-9)
    umh.install(oracle, 1, 1, "du[0] = -p[0] * u[0];", callbacks=cbs)
    _check(g, cpu(oracle, "user", np.ones((1, N)), k[None, :], 0.0, 4.0, **kw))
    late = sa[None, :] > g.t_final[:, None]
    assert np.isnan(g.us[:, 0, :].T[late]).all()


def test_rosenbrock_event_on_robertson(oracle):
    """A continuous event located on the Rosenbrock dense output (Robertson,
    Rodas5): where y3 reaches 1e-3, y2 is halved."""
    N = 800
    say(f"robertson event rodas5 N={N}")
    n, np_, rhs, jac = BUILTIN_SOURCE["robertson"]
    cb = ContinuousCallback("u[2] - 1e-3", affect="u[0] = u[0] + 0.5 * u[1]; u[1] = 0.5 * u[1];")
    w = workloads.robertson(N, seed=114)
    kw = dict(abstol=1e-8, reltol=1e-6, saveat=np.logspace(-3, 4, 15))
    g = run(robertson, w["u0"], w["p"], 0.0, 1e4, alg="rodas5", callback=cb, **kw)
    umh.install(oracle, n, np_, rhs, jac, callbacks=(cb,))
    c = cpu(oracle, "user", w["u0"], w["p"], 0.0, 1e4, alg="rodas5", **kw)
    assert (c["retcode"] == 1).all() and (c["nevent"] == 1).all()
    _check(g, c)
    np.testing.assert_allclose(g.u_final.sum(axis=0), 1.0, atol=1e-10)  # sum y kept


def test_terminate_integrating_backward(oracle):
    """A backward solve (tspan (2, 0)) of u' = -k u from u(2) = 1 grows as
    e^{k (2 - t)}; terminate! where u reaches 2, at t = 2 - ln 2 / k:
    event location on the dense output of negative steps, bit-exact vs the
    oracle."""
    N = 512
    say(f"terminate, backward span N={N}")
    cb = ContinuousCallback("u[0] - 2.0", terminate=True)
    model = UserModel(1, 1, "du[0] = -p[0] * u[0];", callbacks=(cb,))
    rng = np.random.default_rng(115)
    k = rng.uniform(0.5, 3.0, N)
    sa = np.linspace(0.0, 2.0, 9)
    kw = dict(abstol=1e-12, reltol=1e-12, saveat=sa)
    g = run(model, np.ones((1, N)), k[None, :], 2.0, 0.0, **kw)
    assert (g.retcode == 7).all()
    np.testing.assert_allclose(g.t_final, 2.0 - np.log(2.0) / k, rtol=1e-9)
    umh.install(oracle, 1, 1, "du[0] = -p[0] * u[0];", callbacks=(cb,))
    _check(g, cpu(oracle, "user", np.ones((1, N)), k[None, :], 2.0, 0.0, **kw))
    early = sa[None, :] < g.t_final[:, None]  # not reached going backward
    assert np.isnan(g.us[:, 0, :].T[early]).all()


@pytest.mark.parametrize("alg,precision,saves", [("tsit5", "fp64", (True, True)),
                                                 ("rodas5", "fp64", (False, True)),
                                                 ("tsit5", "mixed", (True, False))])
def test_event_points_bitexact(oracle, alg, precision, saves):
    """save_positions: the points saved just before / after each bounce's
    affect, bit-exact against the oracle (times, states, counts; a capacity
    below the count keeps the first points), listed in each solution with
    the saveat points in time order."""
    N = 256
    say(f"event points {alg} {precision} save_positions={saves} N={N}")
    cb = ContinuousCallback("u[0]", affect="u[1] = -p[1] * u[1];", direction=-1,
                            save_positions=saves)
    model = UserModel(2, 2, BALL, callbacks=(cb,))
    rng = np.random.default_rng(114)
    h0 = rng.uniform(5.0, 20.0, N)
    u0 = np.vstack([h0, np.zeros(N)])
    p = np.vstack([np.full(N, 9.81), rng.uniform(0.8, 0.95, N)])
    sa = np.linspace(0.0, 6.0, 7)
    kw = dict(abstol=1e-10, reltol=1e-10, saveat=sa)
    umh.install(oracle, 2, 2, BALL, callbacks=(cb,))
    for cap in (16, 3):
        g = run(model, u0, p, 0.0, 6.0, alg=alg, precision=precision, event_cap=cap, **kw)
        c = cpu(oracle, "user", u0, p, 0.0, 6.0, alg=alg, precision=precision, event_cap=cap,
                **kw)
        _check(g, c)
        per = 2 if saves == (True, True) else 1
        np.testing.assert_array_equal(g.ev_n, per * g.nevent)
        np.testing.assert_array_equal(g.ev_n, c["ev_n"])
        stored = np.arange(cap)[None, :] < np.minimum(g.ev_n, cap)[:, None]
        np.testing.assert_array_equal(np.where(stored, g.ev_t, 0.0), np.where(stored, c["ev_t"], 0.0))
        np.testing.assert_array_equal(np.where(stored[:, :, None], g.ev_u, 0.0),
                                      np.where(stored[:, :, None], c["ev_u"], 0.0))
    # solutions: saveat and event points in time order, saveat first at a tie
    for i in (0, 7, 100):
        sol = g.solution(i)
        assert len(sol.t) == np.count_nonzero(~np.isnan(g.us[:, 0, i])) + min(int(g.ev_n[i]), 3)
        assert (np.diff(sol.t) >= 0).all()


def test_event_points_discrete_and_terminate(oracle):
    """A CallbackSet of a discrete callback that kicks the state at step ends
    inside a window and a terminating continuous callback: the points saved
    before / after each affect (terminate's included) are bit-exact against
    the oracle, and the last one is the final state."""
    N = 512
    say(f"event points, discrete + terminate N={N}")
    cbs = CallbackSet(ContinuousCallback("u[0] - 0.25", terminate=True),
                      DiscreteCallback("t > 0.2 && t < 0.4", affect="u[0] = u[0] * 0.999;"))
    rhs = "du[0] = -p[0] * u[0];"
    model = UserModel(1, 1, rhs, callbacks=cbs)
    rng = np.random.default_rng(115)
    k = rng.uniform(0.5, 3.0, N)
    kw = dict(abstol=1e-10, reltol=1e-10)
    g = run(model, np.ones((1, N)), k[None, :], 0.0, 4.0, event_cap=64, **kw)
    umh.install(oracle, 1, 1, rhs, callbacks=cbs)
    c = cpu(oracle, "user", np.ones((1, N)), k[None, :], 0.0, 4.0, event_cap=64, **kw)
    _check(g, c)
    assert (g.retcode == 7).all() and (g.ev_n == 2 * g.nevent).all()
    np.testing.assert_array_equal(g.ev_n, c["ev_n"])
    np.testing.assert_array_equal(g.ev_t, c["ev_t"])
    np.testing.assert_array_equal(g.ev_u, c["ev_u"])
    for i in (0, 100, N - 1):
        m = int(g.ev_n[i])
        assert g.ev_t[i, m - 1] == g.t_final[i] and g.ev_u[i, m - 1, 0] == g.u_final[0, i]
        sol = g.solution(i)
        # upstream's postamble adds no final point: the last saved one has its time
        assert len(sol.t) == m + 1 and sol.t[0] == 0.0 and sol.t[-1] == g.t_final[i]
        assert sol.u[-1][0] == g.u_final[0, i]


def test_save_everystep_bitexact(oracle):
    """save_everystep: every accepted step's end point, bit-exact against the
    oracle; one point per accepted step, the last at tf with the final state,
    and the solution is (t0, u0) followed by them.  With saveat holding tf the
    last step's point is not repeated (its time was just saved)."""
    N = 300
    say(f"save_everystep lorenz N={N}")
    w = workloads.lorenz(N, seed=116)
    kw = dict(abstol=1e-8, reltol=1e-8, save_everystep=True, event_cap=512)
    src = BUILTIN_SOURCE["lorenz"]
    umh.install(oracle, 3, 3, src[2], jac=src[3], save_everystep=True)
    g = run(lorenz, w["u0"], w["p"], 0.0, 2.0, **kw)
    c = cpu(oracle, "user", w["u0"], w["p"], 0.0, 2.0, event_cap=512, abstol=1e-8,
            reltol=1e-8)
    assert (c["retcode"] == 1).all() and (g.ev_n <= 512).all()
    np.testing.assert_array_equal(g.naccept, c["naccept"])
    np.testing.assert_array_equal(g.u_final, c["u"])
    np.testing.assert_array_equal(g.ev_n, c["ev_n"])
    np.testing.assert_array_equal(g.ev_t, c["ev_t"])
    np.testing.assert_array_equal(g.ev_u, c["ev_u"])
    np.testing.assert_array_equal(g.ev_n, g.naccept)
    for i in (0, 150, N - 1):
        m = int(g.ev_n[i])
        assert (np.diff(g.ev_t[i, :m]) > 0).all() and g.ev_t[i, m - 1] == 2.0
        np.testing.assert_array_equal(g.ev_u[i, m - 1], g.u_final[:, i])
        sol = g.solution(i)
        np.testing.assert_array_equal(sol.t, np.concatenate([[0.0], g.ev_t[i, :m]]))
    sa = np.array([0.7, 2.0])
    g2 = run(lorenz, w["u0"], w["p"], 0.0, 2.0, saveat=sa, **kw)
    c2 = cpu(oracle, "user", w["u0"], w["p"], 0.0, 2.0, saveat=sa, event_cap=512,
             abstol=1e-8, reltol=1e-8)
    np.testing.assert_array_equal(g2.ev_n, c2["ev_n"])
    np.testing.assert_array_equal(g2.ev_t, c2["ev_t"])
    np.testing.assert_array_equal(g2.us, c2["saveat"])
    np.testing.assert_array_equal(g2.ev_n, g2.naccept - 1)
    sol = g2.solution(3)
    assert sol.t[0] == 0.0 and sol.t[-1] == 2.0 and len(sol.t) == int(g2.naccept[3]) + 2


def test_save_everystep_with_callbacks_and_saveat(oracle):
    """save_everystep with a bouncing ball's callback (save_positions true,
    true) and saveat: a bounce step saves its end once before the affect and
    once after it, so each trajectory keeps naccept + nevent points, less the
    last step's (its time tf is a saveat point); all bit-exact against the
    oracle."""
    N = 256
    say(f"save_everystep bouncing ball N={N}")
    cb = ContinuousCallback("u[0]", affect="u[1] = -p[1] * u[1];", direction=-1)
    rng = np.random.default_rng(117)
    u0 = np.vstack([rng.uniform(5.0, 20.0, N), np.zeros(N)])
    p = np.vstack([np.full(N, 9.81), rng.uniform(0.8, 0.95, N)])
    sa = np.linspace(0.0, 6.0, 7)
    kw = dict(abstol=1e-10, reltol=1e-10, saveat=sa, event_cap=1024)
    model = UserModel(2, 2, BALL)
    g = run(model, u0, p, 0.0, 6.0, callback=cb, save_everystep=True, **kw)
    umh.install(oracle, 2, 2, BALL, callbacks=(cb,), save_everystep=True)
    c = cpu(oracle, "user", u0, p, 0.0, 6.0, **kw)
    _check(g, c)
    np.testing.assert_array_equal(g.ev_n, c["ev_n"])
    np.testing.assert_array_equal(g.ev_t, c["ev_t"])
    np.testing.assert_array_equal(g.ev_u, c["ev_u"])
    assert (g.nevent > 0).all()
    # the last step ends at tf = 6, a saveat point: its end is not saved again
    np.testing.assert_array_equal(g.ev_n, g.naccept + g.nevent - 1)


@pytest.mark.parametrize("alg,precision", [("rodas5", "fp64"), ("auto", "mixed")])
def test_save_everystep_rosenbrock(oracle, alg, precision):
    """save_everystep through Rosenbrock steps (Robertson, the built-in
    model's source): step points bit-exact against the oracle, one per
    accepted step, Σy = 1 on every one of them (to 1e-12; 1e-5 with FP32
    stages)."""
    N = 256
    say(f"save_everystep robertson {alg} {precision} N={N}")
    w = workloads.robertson(N, seed=118)
    kw = dict(abstol=1e-8, reltol=1e-6, save_everystep=True, event_cap=1024)
    src = BUILTIN_SOURCE["robertson"]
    umh.install(oracle, 3, 3, src[2], jac=src[3], save_everystep=True)
    g = run(robertson, w["u0"], w["p"], 0.0, 1e3, alg=alg, precision=precision, **kw)
    c = cpu(oracle, "user", w["u0"], w["p"], 0.0, 1e3, alg=alg, precision=precision,
            event_cap=1024, abstol=1e-8, reltol=1e-6)
    assert (c["retcode"] == 1).all() and (g.ev_n <= 1024).all()
    assert_identical(g, c, save=False)
    np.testing.assert_array_equal(g.ev_n, c["ev_n"])
    np.testing.assert_array_equal(g.ev_t, c["ev_t"])
    np.testing.assert_array_equal(g.ev_u, c["ev_u"])
    np.testing.assert_array_equal(g.ev_n, g.naccept)
    stored = ~np.isnan(g.ev_t)
    # FP32 stages: the sum drifts by ~2e-6 (the oracle's mixed mode too)
    np.testing.assert_allclose(g.ev_u.sum(axis=2)[stored], 1.0,
                               atol=1e-5 if precision == "mixed" else 1e-12)


@pytest.mark.parametrize("alg", ["tsit5", "rodas5"])
def test_tstops_dose_bitexact(oracle, alg):
    """tstops with a dosing DiscreteCallback at t == 0.5 (steps land on the
    tstop, the dose applies there): bit-exact against the oracle, the step
    ends (save_everystep) hold the tstops, and the final states follow
    u(1) = (exp(-k/2) + 1) exp(-k/2)."""
    N = 512
    say(f"tstops dose {alg} N={N}")
    rhs = "du[0] = -p[0] * u[0];"
    dose = DiscreteCallback("t == 0.5", affect="u[0] = u[0] + 1.0;")
    rng = np.random.default_rng(119)
    k = rng.uniform(0.2, 5.0, N)
    ts = np.array([0.25, 0.5])
    kw = dict(abstol=1e-10, reltol=1e-10, tstops=ts, save_everystep=True, event_cap=1024)
    g = run(UserModel(1, 1, rhs), np.ones((1, N)), k[None, :], 0.0, 1.0, alg=alg,
            callback=dose, **kw)
    umh.install(oracle, 1, 1, rhs, callbacks=(dose,), save_everystep=True)
    c = cpu(oracle, "user", np.ones((1, N)), k[None, :], 0.0, 1.0, alg=alg, tstops=ts,
            event_cap=1024, abstol=1e-10, reltol=1e-10)
    _check(g, c)
    np.testing.assert_array_equal(g.ev_n, c["ev_n"])
    np.testing.assert_array_equal(g.ev_t, c["ev_t"])
    np.testing.assert_array_equal(g.ev_u, c["ev_u"])
    assert (g.nevent == 1).all() and (g.t_event == 0.5).all()
    for i in (0, 200, N - 1):
        t = g.ev_t[i, :int(g.ev_n[i])]
        assert 0.25 in t and 0.5 in t
    np.testing.assert_allclose(g.u_final[0], (np.exp(-k / 2) + 1.0) * np.exp(-k / 2),
                               rtol=1e-7)
