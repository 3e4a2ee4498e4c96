"""Run-time compiled right-hand sides (hipRTC) on the GPU.  A user model
written like a built-in one reproduces it bit for bit; dual-number
Jacobians, n > 3 and the dense output match the oracle running the same
statements compiled for the host; transcendental functions (device libm vs
host libm) match to tolerance."""
import numpy as np
import pytest

import user_model_host as umh
from gpu_util import ALGS, assert_identical, cpu, gpu, say
from mideq import EnsembleMI355X, MideqError, UserModel, device_solve, workloads
from test_user_model import LORENZ, PENDULUM, ROBERTSON

pytestmark = pytest.mark.gpu

KEYS = ("retcode", "naccept", "nreject", "nf", "nstiff", "u_final", "t_final", "dt_first")
ROB_JAC = ("double y2 = u[1], y3 = u[2], k1 = p[0], k2 = p[1], k3 = p[2];\n"
           "J[0][0] = -k1; J[0][1] = k3 * y3; J[0][2] = k3 * y2;\n"
           "J[1][0] = k1; J[1][1] = -2.0 * k2 * y2 - k3 * y3; J[1][2] = -k3 * y2;\n"
           "J[2][0] = 0.0; J[2][1] = 2.0 * k2 * y2; J[2][2] = 0.0;")
CHAIN5 = ("for (int i = 0; i < 5; i++) {\n"
          "  T left = i > 0 ? u[i - 1] : T(1);\n"
          "  du[i] = p[0] * (left - u[i]) - p[1] * u[i] * u[i] + p[2] * u[(i + 2) % 5];\n"
          "}")
M_LORENZ = UserModel(3, 3, LORENZ, name="lorenz_user")
M_ROB_J = UserModel(3, 3, ROBERTSON, jac=ROB_JAC, name="robertson_user_jac")
M_ROB_AD = UserModel(3, 3, ROBERTSON, name="robertson_user_ad")
M_CHAIN = UserModel(5, 3, CHAIN5, name="chain5")
M_PEND = UserModel(2, 2, PENDULUM, name="pendulum")


def run(f, u0, p, t0, tf, alg="tsit5", precision="fp64", schedule="lane", controller="reference",
        saveat=None, save_dtype="float64", graph=False, **kw):
    N = u0.shape[1]
    t0 = np.broadcast_to(np.asarray(t0, dtype=np.float64), (N,)).copy()
    tf = np.broadcast_to(np.asarray(tf, dtype=np.float64), (N,)).copy()
    ealg = EnsembleMI355X(precision=precision, schedule=schedule, controller=controller,
                          save_dtype=save_dtype, graph=graph)
    return device_solve(f, np.ascontiguousarray(u0), np.ascontiguousarray(p), t0, tf, None,
                        ALGS[alg], ealg, saveat=saveat, **kw)


def _same(a, b):
    for k in KEYS:
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=k)
    if a.us is not None or b.us is not None:
        np.testing.assert_array_equal(a.us, b.us)


@pytest.mark.parametrize("alg,precision,controller,schedule", [
    ("tsit5", "fp64", "reference", "lane"), ("tsit5", "fp64", "fast", "persistent"),
    ("tsit5", "mixed", "fast", "lane"), ("auto", "fp64", "reference", "lane")])
def test_user_lorenz_equals_builtin(alg, precision, controller, schedule):
    N = 3000
    say(f"user lorenz vs built-in {alg} {precision} {controller} {schedule} N={N}")
    w = workloads.lorenz(N, seed=101)
    sa = workloads.saveat_uniform(0.0, 10.0, 21)
    kw = dict(alg=alg, precision=precision, controller=controller, schedule=schedule,
              saveat=sa, abstol=1e-8, reltol=1e-8)
    _same(run(M_LORENZ, w["u0"], w["p"], 0.0, 10.0, **kw),
          gpu("lorenz", w["u0"], w["p"], 0.0, 10.0, **kw))


@pytest.mark.parametrize("alg,precision,controller", [
    ("rodas5", "fp64", "reference"), ("rodas5p", "fp64", "fast"), ("auto", "mixed", "fast")])
def test_user_robertson_with_jacobian_equals_builtin(alg, precision, controller):
    N = 2000
    say(f"user robertson (J given) vs built-in {alg} {precision} {controller} N={N}")
    w = workloads.robertson(N, seed=102)
    sa = np.concatenate([[0.0], np.logspace(-5, 5, 21)])
    kw = dict(alg=alg, precision=precision, controller=controller, saveat=sa, abstol=1e-8,
              reltol=1e-6)
    _same(run(M_ROB_J, w["u0"], w["p"], 0.0, 1e5, **kw),
          gpu("robertson", w["u0"], w["p"], 0.0, 1e5, **kw))


@pytest.mark.parametrize("alg,controller", [("rodas5", "reference"), ("auto", "fast"),
                                            ("rodas5p", "reference")])
def test_user_robertson_dual_jacobian_vs_oracle(oracle, alg, controller):
    N = 1500
    say(f"user robertson (dual-number J) vs oracle {alg} {controller} N={N}")
    umh.install(oracle, 3, 3, ROBERTSON)
    w = workloads.robertson(N, seed=103)
    sa = np.concatenate([[0.0], np.logspace(-5, 5, 21)])
    kw = dict(saveat=sa, abstol=1e-8, reltol=1e-6)
    g = run(M_ROB_AD, w["u0"], w["p"], 0.0, 1e5, alg=alg, controller=controller, **kw)
    c = cpu(oracle, "user", w["u0"], w["p"], 0.0, 1e5, alg=alg,
            controller=1 if controller == "fast" else 0, **kw)
    assert (c["retcode"] == 1).all()
    assert_identical(g, c)


@pytest.mark.parametrize("alg,precision,controller", [
    ("tsit5", "fp64", "fast"), ("tsit5", "mixed", "reference"), ("rodas5", "fp64", "fast"),
    ("auto", "fp64", "reference")])
def test_user_five_states_vs_oracle(oracle, alg, precision, controller):
    """n = 5 (past the built-in models' n <= 3): fast controller's mean(r^2)
    form, reference stiffness test, 5 x 5 register LU."""
    N = 1200
    say(f"user 5-state chain {alg} {precision} {controller} N={N}")
    umh.install(oracle, 5, 3, CHAIN5)
    rng = np.random.default_rng(104)
    u0 = rng.uniform(0.0, 1.0, (5, N))
    p = np.vstack([10.0 ** rng.uniform(-1, 2, N), rng.uniform(0.1, 2.0, N),
                   rng.uniform(0.0, 0.5, N)])
    sa = np.linspace(0.0, 5.0, 11)
    kw = dict(saveat=sa, abstol=1e-8, reltol=1e-6)
    g = run(M_CHAIN, u0, p, 0.0, 5.0, alg=alg, precision=precision, controller=controller, **kw)
    c = cpu(oracle, "user", u0, p, 0.0, 5.0, alg=alg, precision=precision,
            controller=1 if controller == "fast" else 0, **kw)
    assert (c["retcode"] == 1).all()
    assert_identical(g, c)


@pytest.mark.parametrize("alg", ["tsit5", "auto"])
def test_user_transcendental_vs_oracle_to_tolerance(oracle, alg):
    N = 1000
    say(f"user pendulum (libm functions) {alg} N={N}")
    umh.install(oracle, 2, 2, PENDULUM)
    rng = np.random.default_rng(105)
    u0 = np.vstack([rng.uniform(-1, 1, N), rng.uniform(-1, 1, N)])
    p = np.vstack([rng.uniform(0.5, 5.0, N), rng.uniform(0.1, 1.0, N)])
    kw = dict(abstol=1e-10, reltol=1e-10)
    g = run(M_PEND, u0, p, 0.0, 5.0, alg=alg, **kw)
    c = cpu(oracle, "user", u0, p, 0.0, 5.0, alg=alg, **kw)
    assert (g.retcode == 1).all() and (c["retcode"] == 1).all()
    err = np.abs(g.u_final - c["u"]) / (1e-10 + 1e-10 * np.abs(c["u"]))
    assert np.median(err) < 50 and err.max() < 5e3, (np.median(err), err.max())


def test_user_model_graph_replays_and_errors():
    N = 1000
    say(f"user model graph replays, bad source N={N}")
    w = workloads.lorenz(N, seed=106)
    ref = run(M_LORENZ, w["u0"], w["p"], 0.0, 5.0, abstol=1e-8, reltol=1e-8)
    for rep in range(3):
        _same(ref, run(M_LORENZ, w["u0"], w["p"], 0.0, 5.0, graph=True, abstol=1e-8,
                       reltol=1e-8))
    bad = UserModel(2, 1, "du[0] = u[1];\ndu[1] = undefined_fn(u[0]);")
    with pytest.raises(MideqError, match="undefined_fn"):
        run(bad, np.zeros((2, 4)), np.ones((1, 4)), 0.0, 1.0)
    # no parameters: one unused parameter row
    decay = UserModel(1, 0, "du[0] = -T(2) * u[0];")
    r = run(decay, np.ones((1, 8)), np.zeros((1, 8)), 0.0, 1.0, abstol=1e-12, reltol=1e-12)
    np.testing.assert_allclose(r.u_final[0], np.exp(-2.0), rtol=1e-9)


# non-autonomous user models: polynomial forcing, so that device and host
# arithmetic agree bit for bit (no libm calls)
PR_POLY = "du[0] = p[0] * (u[0] - (t * t - t)) + (2.0 * t - 1.0);"
FORCED = "du[0] = u[1];\ndu[1] = -p[0] * p[0] * u[0] + t;"


@pytest.mark.parametrize("alg,precision,controller", [
    ("rodas5", "fp64", "reference"), ("rodas5p", "fp64", "fast"), ("rodas5", "mixed", "reference"),
    ("auto", "fp64", "reference"), ("tsit5", "fp64", "fast")])
def test_user_time_dependent_prothero_vs_oracle(oracle, alg, precision, controller):
    """y' = lambda (y - phi(t)) + phi'(t), phi = t^2 - t, y(0) = phi(0): the
    solution is phi.  Stage times, the d_i dt df/dt terms and the dual-number
    df/dt bit-exact against the oracle, saveat included."""
    N = 1024
    say(f"user non-autonomous Prothero-Robinson {alg} {precision} {controller} N={N}")
    umh.install(oracle, 1, 1, PR_POLY)
    model = UserModel(1, 1, PR_POLY, name="prothero_poly")
    rng = np.random.default_rng(107)
    lam = -(10.0 ** rng.uniform(0.0, 4.0 if alg != "tsit5" else 1.0, N))
    sa = np.linspace(0.0, 2.0, 9)
    kw = dict(saveat=sa, abstol=1e-8, reltol=1e-6)
    g = run(model, np.zeros((1, N)), lam[None, :], 0.0, 2.0, alg=alg, precision=precision,
            controller=controller, **kw)
    c = cpu(oracle, "user", np.zeros((1, N)), lam[None, :], 0.0, 2.0, alg=alg,
            precision=precision, controller=1 if controller == "fast" else 0, **kw)
    assert (c["retcode"] == 1).all()
    assert_identical(g, c)
    if precision == "fp64":
        np.testing.assert_allclose(g.u_final[0], 2.0, atol=1e-5)


@pytest.mark.parametrize("alg", ["tsit5", "auto5p"])
def test_user_time_dependent_forced_oscillator(oracle, alg):
    """x'' = -w^2 x + t from rest, saveat on [0, 6]: bit-exact against the
    oracle and within tolerance of x = t / w^2 - sin(w t) / w^3."""
    N = 2048
    say(f"user forced oscillator {alg} N={N}")
    umh.install(oracle, 2, 1, FORCED)
    model = UserModel(2, 1, FORCED, name="forced_oscillator")
    w = np.linspace(0.5, 3.0, N)
    sa = np.linspace(0.0, 6.0, 13)
    kw = dict(saveat=sa, abstol=1e-10, reltol=1e-10)
    g = run(model, np.zeros((2, N)), w[None, :], 0.0, 6.0, alg=alg, **kw)
    c = cpu(oracle, "user", np.zeros((2, N)), w[None, :], 0.0, 6.0, alg=alg, **kw)
    assert (c["retcode"] == 1).all()
    assert_identical(g, c)
    x = sa[:, None] / w[None, :] ** 2 - np.sin(sa[:, None] * w[None, :]) / w[None, :] ** 3
    np.testing.assert_allclose(g.us[:, 0, :], x, rtol=1e-6, atol=1e-8)


CHAIN8 = ("for (int i = 0; i < 8; i++) {\n"
          "  T left = i > 0 ? u[i - 1] : T(1);\n"
          "  du[i] = p[0] * (left - u[i]) - p[1] * u[i] * u[i] + p[2] * u[(i + 3) % 8];\n"
          "}")


@pytest.mark.parametrize("alg,precision", [("tsit5", "fp64"), ("rodas5", "fp64"),
                                           ("auto", "mixed")])
def test_user_eight_states(oracle, alg, precision):
    """n = 8, the largest model whose dual Jacobian takes all directions in
    one pass (8 x 8 register LU, dual numbers with eight directions), with
    saveat: bit-exact against the oracle."""
    N = 700
    say(f"user 8-state chain {alg} {precision} N={N}")
    umh.install(oracle, 8, 3, CHAIN8)
    model = UserModel(8, 3, CHAIN8, name="chain8")
    rng = np.random.default_rng(108)
    u0 = rng.uniform(0.0, 1.0, (8, N))
    p = np.vstack([10.0 ** rng.uniform(-1, 2, N), rng.uniform(0.1, 2.0, N),
                   rng.uniform(0.0, 0.5, N)])
    kw = dict(saveat=np.linspace(0.0, 3.0, 7), abstol=1e-8, reltol=1e-6)
    g = run(model, u0, p, 0.0, 3.0, alg=alg, precision=precision, **kw)
    c = cpu(oracle, "user", u0, p, 0.0, 3.0, alg=alg, precision=precision, **kw)
    assert (c["retcode"] == 1).all()
    assert_identical(g, c)


L96 = ("for (int i = 0; i < 16; i++)\n"
       "  du[i] = (u[(i + 1) % 16] - u[(i + 14) % 16]) * u[(i + 15) % 16]"
       " - p[1 + i % 2] * u[i] + p[0];")


@pytest.mark.parametrize("alg,precision", [("tsit5", "fp64"), ("rodas5", "fp64"),
                                           ("auto", "mixed")])
def test_user_sixteen_states_at_the_limit(oracle, alg, precision):
    """n = 16, the largest one-lane model: a Lorenz-96 ring whose odd sites
    are damped up to 2000 (stiff), with saveat.  Its Jacobian comes from
    dual numbers one direction at a time (n > 8); the oracle's host build
    differentiates with all 16 directions at once.  Bit-exact."""
    N = 500
    say(f"user 16-state Lorenz-96 {alg} {precision} N={N}")
    umh.install(oracle, 16, 3, L96)
    model = UserModel(16, 3, L96, name="lorenz96_16")
    rng = np.random.default_rng(116)
    u0 = rng.uniform(-1.0, 1.0, (16, N))
    p = np.vstack([rng.uniform(2.0, 8.0, N), rng.uniform(0.5, 1.5, N),
                   10.0 ** rng.uniform(0, 3.3, N)])
    kw = dict(saveat=np.linspace(0.0, 2.0, 5), abstol=1e-8, reltol=1e-6)
    g = run(model, u0, p, 0.0, 2.0, alg=alg, precision=precision, **kw)
    c = cpu(oracle, "user", u0, p, 0.0, 2.0, alg=alg, precision=precision, **kw)
    assert (c["retcode"] == 1).all()
    assert_identical(g, c)
