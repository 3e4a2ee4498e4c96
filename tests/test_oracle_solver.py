"""Pin the oracle's integrator against closed forms, invariants, independent
solvers and the reference's own known-answer tests (test/core.jl)."""
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp
from scipy.linalg import expm

LORENZ_P = [10.0, 28.0, 8.0 / 3.0]
ROBER_P = [0.04, 3.0e7, 1.0e4]


def lorenz_f(t, u, s, r, b):
    return [s * (u[1] - u[0]), u[0] * (r - u[2]) - u[1], u[0] * u[1] - b * u[2]]


def rober_f(t, y, k1, k2, k3):
    return [-k1 * y[0] + k3 * y[1] * y[2], k1 * y[0] - k2 * y[1] ** 2 - k3 * y[1] * y[2],
            k2 * y[1] ** 2]


def rober_j(t, y, k1, k2, k3):
    return [[-k1, k3 * y[2], k3 * y[1]], [k1, -2 * k2 * y[1] - k3 * y[2], -k3 * y[1]],
            [0, 2 * k2 * y[1], 0]]


# ----------------------------------------------------------- core.jl known answers
def test_core_basic_ode_solve(oracle):
    """test/core.jl:9-19: u'=1.01u, u0=0.5 on (0,1), Tsit5 at 1e-8."""
    r = oracle.solve("exp", [0.5], [1.01], (0.0, 1.0), abstol=1e-8, reltol=1e-8)
    assert r["retcode"] == 1
    exact = 0.5 * math.exp(1.01)
    assert abs(r["u"][0] - exact) / exact < 2e-8
    assert r["t"] == 1.0


def test_core_saveat_eleven_points(oracle):
    """test/core.jl:90-96: saveat=0.1 on (0,1) gives 11 points, both ends."""
    sa = np.array([i / 10 for i in range(11)])
    r = oracle.solve("exp", [0.5], [1.01], (0.0, 1.0), saveat=sa)
    s = r["saveat"][:, 0]
    assert len(s) == 11 and np.all(np.isfinite(s))
    assert s[0] == 0.5  # exact copy of u0 (core.jl:34 asserts ==)
    assert s[-1] == r["u"][0]
    exact = 0.5 * np.exp(1.01 * sa)
    # default tolerances abstol=1e-6, reltol=1e-3
    assert np.max(np.abs(s - exact) / exact) < 2e-3


def test_core_inplace_lorenz_and_dense(oracle):
    """test/core.jl:21-36 and 51-58."""
    r = oracle.solve("lorenz", [1.0, 0.0, 0.0], LORENZ_P, (0.0, 1.0),
                     saveat=np.array([0.0, 0.5, 1.0]))
    assert r["retcode"] == 1 and len(r["u"]) == 3
    np.testing.assert_array_equal(r["saveat"][0], [1.0, 0.0, 0.0])
    r = oracle.solve("exp", [0.5], [1.01], (0.0, 1.0), saveat=np.array([0.5]))
    assert r["saveat"][0, 0] > 0.5


def test_core_robertson_success(oracle):
    """test/core.jl:38-49 (Rodas5 stands in for Rodas5P)."""
    r = oracle.solve("robertson", [1.0, 0.0, 0.0], ROBER_P, (0.0, 1e5), alg="rodas5")
    assert r["retcode"] == 1


# ----------------------------------------------------------- initdt
def test_initdt_hand_derived(oracle):
    """Hairer-Norsett-Wanner II.4 for u'=1.01u, u0=0.5, tol 1e-8, order 5."""
    u0, lam, tol = 0.5, 1.01, 1e-8
    sk = tol + u0 * tol
    d0 = u0 / sk
    f0 = lam * u0
    d1 = f0 / sk
    dt0 = (d0 / d1) / 100
    u1 = u0 + dt0 * f0
    d2 = abs(lam * u1 - f0) / sk / dt0
    dt1 = 10 ** (-(2 + math.log10(max(d1, d2))) / 5)
    expect = min(100 * dt0, dt1, 1.0)
    got = oracle.initdt("exp", [u0], [lam], (0.0, 1.0), abstol=tol, reltol=tol)
    assert got == pytest.approx(expect, rel=1e-10)
    r = oracle.solve("exp", [u0], [lam], (0.0, 1.0), abstol=tol, reltol=tol)
    assert r["dt0"] == got


def test_initdt_small_norm_branch(oracle):
    """d0 < 1e-5 (u0 = 0) takes the 1e-6 branch, then the f1==f0 early return."""
    got = oracle.initdt("exp", [0.0], [1.0], (0.0, 1.0), abstol=1e-8, reltol=1e-8)
    assert got == pytest.approx(100 * 1e-6)


# ----------------------------------------------------------- controller
def _pi_expect(E, qold, dt, qsteady_max):
    b1, b2, gam, qmin, qmax = 7 / 50, 2 / 25, 0.9, 0.2, 10.0
    if E == 0:
        q, q11 = 1 / qmax, None
    else:
        q11 = E ** b1
        q = max(1 / qmax, min(1 / qmin, q11 / qold ** b2 / gam))
    if E <= 1:
        if 1 <= q <= qsteady_max:
            q = 1.0
        return True, dt / q, max(E, 1e-4)
    return False, dt / min(1 / qmin, q11 / gam), qold


@pytest.mark.parametrize("E,qold,qs", [(0.0, 1e-4, 1.0), (0.3, 1e-4, 1.0), (0.3, 0.5, 1.0),
                                       (0.99, 0.9, 1.2), (1.0, 0.2, 1.0), (1.5, 0.3, 1.0),
                                       (1e6, 0.3, 1.0), (0.5, 0.5, 1.2), (2e-9, 1e-4, 1.0)])
def test_pi_controller(oracle, E, qold, qs):
    acc, dtn, qn = oracle.controller(E, qold, 0.1, qs)
    eacc, edtn, eqn = _pi_expect(E, qold, 0.1, qs)
    assert acc == eacc
    assert dtn == pytest.approx(edtn, rel=1e-10)
    assert qn == eqn


def test_ctl_pow_accuracy(oracle):
    rng = np.random.default_rng(1)
    for x in np.concatenate([10 ** rng.uniform(-12, 8, 400), [1.0, 0.5, 2.0, 1e-4]]):
        for y in (0.14, 0.08, -0.2, 1 / 5):
            assert oracle.ctl_pow(x, y) == pytest.approx(x ** y, rel=5e-11)
    assert oracle.ctl_exp2(3.0) == 8.0 and oracle.ctl_log2(1024.0) == 10.0


def test_fast_log2_accuracy(oracle):
    """Division-free log2 of the fast controller: |error| < 1.4e-8 from its
    degree-8 fit, exact at powers of two, IEEE special values."""
    rng = np.random.default_rng(2)
    xs = np.concatenate([10 ** rng.uniform(-300, 300, 3000), rng.uniform(0.5, 2.0, 3000)])
    err = max(abs(oracle.fast_log2(float(x)) - math.log2(x)) for x in xs)
    assert err < 1.4e-8, err
    for k in (-1000, -3, 0, 1, 7, 1000):
        assert oracle.fast_log2(2.0 ** k) == float(k)
    assert oracle.fast_log2(0.0) == -math.inf and oracle.fast_log2(math.inf) == math.inf
    assert math.isnan(oracle.fast_log2(-1.0))


def test_fast_exp2_accuracy(oracle):
    """2^z on the fast controller's clamped domain |z| <= log2(10): relative
    error < 2e-7 (degree-6 series), exact at integers."""
    zs = np.linspace(-math.log2(10.0), math.log2(10.0), 4001)
    err = max(abs(oracle.fast_exp2(float(z)) / 2.0 ** z - 1.0) for z in zs)
    assert err < 2e-7, err
    for k in range(-3, 4):
        assert oracle.fast_exp2(float(k)) == 2.0 ** k


@pytest.mark.parametrize("model,alg,tol", [("lorenz", "tsit5", (1e-8, 1e-8)),
                                           ("robertson", "rodas5", (1e-8, 1e-6)),
                                           ("vdp", "auto", (1e-8, 1e-6))])
def test_fast_controller_tracks_reference_controller(oracle, model, alg, tol):
    """The fast controller is the same PI law evaluated differently: step
    counts within 2 % and final-state errors of the same size as the
    reference controller's, against a tight solve."""
    w = {"lorenz": ([1.0, 0.0, 0.0], LORENZ_P, (0.0, 10.0)),
         "robertson": ([1.0, 0.0, 0.0], ROBER_P, (0.0, 1e5)),
         "vdp": ([2.0, 0.0], [300.0], (0.0, 600.0))}[model]
    u0, p, span = w
    at, rt = tol
    ref = oracle.solve(model, u0, p, span, alg=alg, abstol=at * 1e-4, reltol=rt * 1e-4)
    a = oracle.solve(model, u0, p, span, alg=alg, abstol=at, reltol=rt, controller=0)
    b = oracle.solve(model, u0, p, span, alg=alg, abstol=at, reltol=rt, controller=1)
    assert a["retcode"] == 1 and b["retcode"] == 1
    na, nb = a["naccept"] + a["nreject"], b["naccept"] + b["nreject"]
    assert abs(na - nb) <= 0.02 * na + 2, (na, nb)
    ea = np.max(np.abs(a["u"] - ref["u"]) / (at + rt * np.abs(ref["u"])))
    eb = np.max(np.abs(b["u"] - ref["u"]) / (at + rt * np.abs(ref["u"])))
    assert eb < 3.0 * ea + 1.0, (ea, eb)


# ----------------------------------------------------------- Tsit5 step vs linear theory
def test_tsit5_step_linear_stability_polynomial(oracle):
    """On u'=lam u one step is u1 = R(z) u0 and the embedded error is
    u0 (Rhat(z) - R(z)); R, Rhat from the tableau by linear algebra."""
    c, A, bt, R_ = oracle.tsit5_tableau()
    b = A[6]
    bhat = b + bt
    one = np.ones(7)
    for lam, dt in ((1.01, 0.1), (-3.0, 0.2), (0.5, 1.0)):
        z = lam * dt
        M = np.linalg.solve(np.eye(7) - z * A, one)
        Rz = 1 + z * b @ M
        Rh = 1 + z * bhat @ M
        u0 = 0.7
        u1, E, k = oracle.tsit5_step("exp", [u0], [lam], 0.0, dt, 1e-6, 1e-3)
        assert u1[0] == pytest.approx(Rz * u0, rel=1e-14)
        sc = 1e-6 + max(abs(u0), abs(u1[0])) * 1e-3
        assert E == pytest.approx(abs(u0 * (Rh - Rz)) / sc, rel=1e-9)


def test_tsit5_convergence_order(oracle):
    """Fixed-step global error on a linear 3x3 system vs expm: ratio 2^5."""
    M = np.array([[-0.5, 1.0, 0.2], [-1.0, -0.3, 0.0], [0.1, 0.4, -0.8]])
    u0 = np.array([1.0, -0.5, 2.0])
    T = 2.0
    exact = expm(M * T) @ u0

    def run(N):
        u = u0.copy()
        for i in range(N):
            u, _, _ = oracle.tsit5_step("linear", u, M.reshape(-1), i * T / N, T / N, 1, 1)
        return np.linalg.norm(u - exact)

    e = [run(N) for N in (20, 40, 80)]
    rates = [math.log2(e[i] / e[i + 1]) for i in range(2)]
    assert all(4.6 < r < 5.5 for r in rates), rates


def test_tsit5_interpolant_accuracy(oracle):
    """Dense output on u'=lam u: theta=0 and 1 reproduce the endpoints, local
    error of the 4th-order interpolant scales like dt^5."""
    lam, u0 = -1.3, 1.0
    errs = []
    for dt in (0.2, 0.1, 0.05):
        u1, _, k = oracle.tsit5_step("exp", [u0], [lam], 0.0, dt, 1, 1)
        assert oracle.tsit5_interp([u0], k, dt, 0.0)[0] == u0
        assert oracle.tsit5_interp([u0], k, dt, 1.0)[0] == pytest.approx(u1[0], abs=1e-15)
        th = 0.4
        errs.append(abs(oracle.tsit5_interp([u0], k, dt, th)[0] - u0 * math.exp(lam * th * dt)))
    assert 4.5 < math.log2(errs[0] / errs[1]) < 5.6
    assert 4.5 < math.log2(errs[1] / errs[2]) < 5.6


# ----------------------------------------------------------- Rodas5
@pytest.mark.parametrize("which", [0, 1], ids=["rodas5", "rodas5p"])
def test_rodas5_convergence_order(oracle, which):
    M = np.array([[-0.5, 1.0, 0.2], [-1.0, -0.3, 0.0], [0.1, 0.4, -0.8]])
    u0 = np.array([1.0, -0.5, 2.0])
    T = 2.0
    exact = expm(M * T) @ u0

    def run(N):
        u = u0.copy()
        for i in range(N):
            rc, u, _ = oracle.rodas5_step("linear", u, M.reshape(-1), i * T / N, T / N, 1, 1,
                                          which)
            assert rc == 0
        return np.linalg.norm(u - exact)

    e = [run(N) for N in (8, 16, 32)]
    rates = [math.log2(e[i] / e[i + 1]) for i in range(2)]
    # Rodas5P's stability function matches exp(z) through z^6 (see the
    # tableau tests), so on a linear problem it converges at order 6
    lo, hi = (4.6, 5.6) if which == 0 else (5.5, 6.6)
    assert all(lo < r < hi for r in rates), rates


@pytest.mark.parametrize("which", [0, 1], ids=["rodas5", "rodas5p"])
def test_rodas5_nonlinear_order(oracle, which):
    """Nonlinear (Lorenz, short horizon): fixed-step order 5 against a tight
    DOP853 solution; exercises the multi-child (alpha) order conditions."""
    u0 = np.array([1.0, 1.0, 1.0])
    T = 0.2
    ref = solve_ivp(lorenz_f, (0, T), u0, method="DOP853", rtol=1e-14, atol=1e-14,
                    args=tuple(LORENZ_P)).y[:, -1]

    def run(N):
        u = u0.copy()
        for i in range(N):
            rc, u, _ = oracle.rodas5_step("lorenz", u, LORENZ_P, i * T / N, T / N, 1, 1, which)
        return np.linalg.norm(u - ref)

    e = [run(N) for N in (10, 20, 40)]
    rates = [math.log2(e[i] / e[i + 1]) for i in range(2)]
    assert all(4.5 < r < 5.8 for r in rates), rates


@pytest.mark.parametrize("alg", ["rodas5", "rodas5p"])
def test_rodas5_prothero_robinson(oracle, alg):
    """y' = lam (y - sin t) + cos t, y(0)=0, exact sin t; stiff lam, large
    steps: a stiffly accurate L-stable method stays accurate (the
    non-autonomous d_i*dt*dT term and c_i are exercised here)."""
    for lam in (-1e2, -1e6):
        r = oracle.solve("prothero", [0.0], [lam], (0.0, 2.0), alg=alg, abstol=1e-9,
                         reltol=1e-9)
        assert r["retcode"] == 1
        assert abs(r["u"][0] - math.sin(2.0)) < 5e-8, (lam, r)


@pytest.mark.parametrize("which,g", [(0, 0.19), (1, 0.21193756319429014)],
                         ids=["rodas5", "rodas5p"])
def test_rodas5_singular_W_rejected(oracle, which, g):
    # W = J - I/(gamma dt) singular when lam = 1/(gamma dt): EEst := 2
    dt = 0.5
    rc, u1, E = oracle.rodas5_step("exp", [1.0], [1.0 / (g * dt)], 0.0, dt, 1e-6, 1e-3, which)
    assert rc == -1 and E == 2.0


@pytest.mark.parametrize("alg", ["rodas5", "rodas5p"])
def test_robertson_invariant_and_radau(oracle, alg):
    """Sum y = 1 is preserved (linear invariant); final state vs Radau."""
    r = oracle.solve("robertson", [1.0, 0.0, 0.0], ROBER_P, (0.0, 1e5), alg=alg,
                     abstol=1e-8, reltol=1e-6)
    assert r["retcode"] == 1
    assert abs(r["u"].sum() - 1.0) < 1e-13
    ref = solve_ivp(rober_f, (0, 1e5), [1, 0, 0], method="Radau", rtol=1e-12, atol=1e-14,
                    jac=rober_j, args=tuple(ROBER_P)).y[:, -1]
    w = np.abs(r["u"] - ref) / (1e-8 + 1e-6 * np.abs(ref))
    assert w.max() < 10.0, (r["u"], ref)
    # Rodas5 at rtol 1e-6 takes O(100) steps (Radau 166, BDF 288 in a scipy probe)
    assert 50 < r["naccept"] < 400


# ----------------------------------------------------------- independent solvers
def test_lorenz_vs_dop853(oracle):
    ref = solve_ivp(lorenz_f, (0, 10), [1, 0, 0], method="DOP853", rtol=1e-13, atol=1e-13,
                    args=tuple(LORENZ_P)).y[:, -1]
    r = oracle.solve("lorenz", [1, 0, 0], LORENZ_P, (0.0, 10.0), abstol=1e-12, reltol=1e-12)
    assert np.linalg.norm(r["u"] - ref) / np.linalg.norm(ref) < 1e-9
    r8 = oracle.solve("lorenz", [1, 0, 0], LORENZ_P, (0.0, 10.0), abstol=1e-8, reltol=1e-8)
    err8 = np.linalg.norm(r8["u"] - ref) / np.linalg.norm(ref)
    assert 1e-10 < err8 < 1e-5
    assert 600 < r8["naccept"] < 1100


def test_vdp_classic_vs_radau(oracle):
    mu = 10.0

    def f(t, u):
        return [u[1], mu * (1 - u[0] ** 2) * u[1] - u[0]]

    ref = solve_ivp(f, (0, 2 * mu), [2, 0], method="Radau", rtol=1e-12, atol=1e-12).y[:, -1]
    for alg in ("tsit5", "rodas5", "auto", "rodas5p", "auto5p"):
        r = oracle.solve("vdp", [2.0, 0.0], [mu], (0.0, 2 * mu), alg=alg, abstol=1e-10,
                         reltol=1e-10)
        assert r["retcode"] == 1
        assert np.abs(r["u"] - ref).max() < 1e-5 * (1 + np.abs(ref).max()), (alg, r["u"], ref)


def test_brusselator_vs_bdf(oracle):
    N = 8
    A, B, alpha = 1.0, 3.0, 0.02
    x = np.arange(N) / N
    u0 = np.empty(2 * N)
    u0[0::2] = 1 + np.sin(2 * np.pi * x)
    u0[1::2] = 3.0
    p = [A, B, alpha]

    def f(t, y):
        return oracle.rhs("brusselator", y, p)

    def jac(t, y):
        return oracle.jac("brusselator", y, p)

    # Jacobian matches finite differences of the RHS
    J = jac(0, u0)
    eps = 1e-7
    Jfd = np.stack([(f(0, u0 + eps * e) - f(0, u0 - eps * e)) / (2 * eps)
                    for e in np.eye(2 * N)], axis=1)
    np.testing.assert_allclose(J, Jfd, atol=1e-6)
    ref = solve_ivp(f, (0, 10), u0, method="BDF", jac=jac, rtol=1e-11, atol=1e-11).y[:, -1]
    for alg in ("tsit5", "rodas5", "rodas5p"):
        r = oracle.solve("brusselator", u0, p, (0.0, 10.0), alg=alg, abstol=1e-9, reltol=1e-9)
        assert r["retcode"] == 1
        assert np.abs(r["u"] - ref).max() < 1e-6, alg


def test_model_jacobians_fd(oracle):
    rng = np.random.default_rng(3)
    cases = [("lorenz", 3, LORENZ_P), ("robertson", 3, [0.04, 3e7, 1e4]), ("vdp", 2, [7.0]),
             ("exp", 1, [1.5])]
    for model, n, p in cases:
        u = rng.uniform(0.1, 1.0, n) * (1e-3 if model == "robertson" else 1)
        J = oracle.jac(model, u, p)
        h = 1e-6 * np.maximum(1e-6, np.abs(u))
        Jfd = np.stack([(oracle.rhs(model, u + h[i] * e, p) - oracle.rhs(model, u - h[i] * e, p))
                        / (2 * h[i]) for i, e in enumerate(np.eye(n))], axis=1)
        np.testing.assert_allclose(J, Jfd, rtol=1e-5, atol=1e-6 * np.abs(J).max())


# ----------------------------------------------------------- LU
@pytest.mark.parametrize("n", [1, 2, 3, 7, 16, 64, 128])
def test_lu_vs_numpy(oracle, n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    b = rng.standard_normal(n)
    rc, x = oracle.lu_solve(A, b)
    assert rc == 0
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-9)


def test_lu_pivoting_and_singular(oracle):
    A = np.array([[0.0, 1.0], [1.0, 0.0]])  # needs a row swap
    rc, x = oracle.lu_solve(A, np.array([2.0, 3.0]))
    assert rc == 0 and list(x) == [3.0, 2.0]
    rc, _ = oracle.lu_solve(np.array([[1.0, 2.0], [2.0, 4.0]]), np.array([1.0, 1.0]))
    assert rc == -1


# ----------------------------------------------------------- tstops, retcodes
@pytest.mark.parametrize("tf", [0.3, 1.0 / 3.0, 10.0, 7.77])
def test_final_time_is_exactly_tf(oracle, tf):
    r = oracle.solve("lorenz", [1, 0, 0], LORENZ_P, (0.0, tf), abstol=1e-6, reltol=1e-6)
    assert r["t"] == tf and r["retcode"] == 1


def test_retcodes(oracle):
    r = oracle.solve("lorenz", [1, 0, 0], LORENZ_P, (0.0, 10.0), maxiters=5)
    assert r["retcode"] == 2 and r["naccept"] + r["nreject"] == 5
    r = oracle.solve("lorenz", [1, 0, 0], LORENZ_P, (0.0, 10.0), dtmin=0.5)
    assert r["retcode"] == 3
    r = oracle.solve("lorenz", [1, 0, 0], [float("nan"), 28, 8 / 3], (0.0, 1.0))
    assert r["retcode"] in (4, 5)
    r = oracle.solve("exp", [1.0], [1.0], (0.0, 0.0))
    assert r["retcode"] == 1 and r["naccept"] == 0
    # a NaN end is not integrated (InitialFailure), never reported as Success
    r = oracle.solve("exp", [1.0], [1.0], (1.0, float("nan")))
    assert r["retcode"] == 6 and r["naccept"] == 0 and r["u"][0] == 1.0


@pytest.mark.parametrize("alg", ["tsit5", "rodas5", "rodas5p", "auto"])
def test_backward_integration_closed_form(oracle, alg):
    """tspan (1, 0): integrated backward with negative steps, as upstream
    does for tf < t0.  u' = 1.01 u from u(1) = e^1.01 gives u(0) = 1
    (test/core.jl:10-14 run in reverse); save points in ascending order are
    filled from the last one down, points beyond t0 stay NaN."""
    sa = np.array([-0.5, 0.0, 0.25, 0.5, 1.0, 1.5])
    r = oracle.solve("exp", [np.exp(1.01)], [1.01], (1.0, 0.0), alg=alg, abstol=1e-12,
                     reltol=1e-12, saveat=sa)
    assert r["retcode"] == 1 and r["t"] == 0.0 and r["naccept"] > 0
    assert abs(r["u"][0] - 1.0) < 1e-9
    s = r["saveat"][:, 0]
    assert np.isnan(s[0]) and np.isnan(s[5])  # before tf / beyond t0
    np.testing.assert_allclose(s[1:4], np.exp(1.01 * sa[1:4]), rtol=1e-9)
    assert s[4] == np.exp(1.01)  # the point at t0 is u0 itself


def test_backward_then_forward_lorenz(oracle):
    """Lorenz over (0, 1) forward, then back from the end state over (1, 0):
    the start comes back (to 1e-6: the reversed flow amplifies the error)."""
    f = oracle.solve("lorenz", [1, 0, 0], LORENZ_P, (0.0, 1.0), abstol=1e-12, reltol=1e-12)
    b = oracle.solve("lorenz", f["u"], LORENZ_P, (1.0, 0.0), abstol=1e-12, reltol=1e-12)
    assert b["retcode"] == 1 and b["t"] == 0.0
    np.testing.assert_allclose(b["u"], [1.0, 0.0, 0.0], atol=1e-6)


# ----------------------------------------------------------- ensemble
def test_ensemble_threads_matches_serial(oracle):
    rng = np.random.default_rng(7)
    N = 64
    p = np.array(LORENZ_P)[:, None] * (1 + rng.uniform(-0.1, 0.1, (3, N)))
    u0 = np.zeros((3, N))
    u0[0] = 1.0
    sa = np.linspace(0, 2, 7)
    a = oracle.ensemble("lorenz", u0, p, 0.0, 2.0, nthreads=1, abstol=1e-8, reltol=1e-8,
                        saveat=sa)
    b = oracle.ensemble("lorenz", u0, p, 0.0, 2.0, nthreads=4, abstol=1e-8, reltol=1e-8,
                        saveat=sa)
    for k in ("u", "saveat", "naccept", "nreject", "retcode"):
        np.testing.assert_array_equal(a[k], b[k])
    for i in (0, 17, 63):
        r = oracle.solve("lorenz", u0[:, i], p[:, i], (0.0, 2.0), abstol=1e-8, reltol=1e-8,
                         saveat=sa)
        np.testing.assert_array_equal(r["u"], a["u"][:, i])
        np.testing.assert_array_equal(r["saveat"], a["saveat"][:, :, i])


# ----------------------------------------------------------- mixed precision
def test_mixed_precision_close_to_fp64(oracle):
    r64 = oracle.solve("lorenz", [1, 0, 0], LORENZ_P, (0.0, 10.0), abstol=1e-6, reltol=1e-6)
    r32 = oracle.solve("lorenz", [1, 0, 0], LORENZ_P, (0.0, 10.0), abstol=1e-6, reltol=1e-6,
                       precision="mixed")
    assert r32["retcode"] == 1 and r32["t"] == 10.0
    rel = np.linalg.norm(r32["u"] - r64["u"]) / np.linalg.norm(r64["u"])
    assert rel < 1e-3
    rr = oracle.solve("robertson", [1, 0, 0], ROBER_P, (0.0, 1e5), alg="rodas5", abstol=1e-8,
                      reltol=1e-6, precision="mixed")
    # FP32 RHS rounding does not conserve sum y exactly; drift stays ~reltol
    assert rr["retcode"] == 1 and abs(rr["u"].sum() - 1) < 1e-5


# ----------------------------------------------------------- auto-switching
def test_autoswitch(oracle):
    r = oracle.solve("lorenz", [1, 0, 0], LORENZ_P, (0.0, 10.0), alg="auto", abstol=1e-8,
                     reltol=1e-8)
    assert r["nswitch"] == 0 and r["nstiff"] == 0
    r = oracle.solve("robertson", [1, 0, 0], ROBER_P, (0.0, 1e5), alg="auto", abstol=1e-8,
                     reltol=1e-6)
    assert r["retcode"] == 1 and r["nswitch"] >= 1 and r["first_stiff_iter"] > 10
    assert r["nstiff"] > 0.5 * r["naccept"]
    assert abs(r["u"].sum() - 1) < 1e-12


# --------------------------------------------- AutoSwitch pinned by closed forms
# AutoTsit5's switching rule (OrdinaryDiffEq AutoSwitch): before every attempt
# but the first, the attempt's dt times the last stiffness estimate over
# Tsit5's real stability size 3.5068 is compared with 9/10; consecutive
# positive results count up, negative ones down; more than 10 positives in
# Tsit5 switch to the Rosenbrock method with dt * 2, more than 3 negatives in
# the Rosenbrock method switch back with dt / 2.  These values are written
# here from that definition, not read from the oracle.
AS_STAB, AS_TOL, AS_MAXSTIFF, AS_MAXNONSTIFF, AS_DTFAC = 3.5068, 0.9, 10, 3, 2.0

# u' = lambda(t) (u - g(t)) + g'(t), g = 1 / (1 + t^2) (the solution from
# u(0) = 1), |lambda| dropping from p0 to p1 past t = p2 (lambda rational in
# t, so host and device evaluate it identically): stiff, then not
DROP_LAM = ("T s = t / p[2]; T s2 = s * s; T s4 = s2 * s2; T s8 = s4 * s4;\n"
            "T lam = -(p[1] + (p[0] - p[1]) / (1.0 + s8));\n")
DROP_RHS = DROP_LAM + "T q = 1.0 / (1.0 + t * t);\ndu[0] = lam * (u[0] - q) - 2.0 * t * q * q;"
DROP_JAC = DROP_LAM.replace("T ", "double ") + "J[0][0] = lam;"


DROP_PARAMS = ([1e4, 1.0, 1.0], [3e3, 0.5, 1.5], [2e4, 2.0, 0.8])


def lam_drop(p, t):
    s = t / p[2]
    s2 = s * s
    s4 = s2 * s2
    s8 = s4 * s4
    return -(p[1] + (p[0] - p[1]) / (1.0 + s8))


def replay_autoswitch(tr, eig_exact, tf, dtmax, start_mode=0):
    """Replays the switching rule on a traced solve with the exact stiffness
    estimate of every attempt (eig_exact(record)); asserts the method and the
    dt of every attempt.  Returns (switches to stiff, switches back) or None
    when a test lands within 1e-6 of the threshold (undecidable here)."""
    count, mode, up, down = 0, start_mode, 0, 0
    for i, r in enumerate(tr):
        dt = r["dt_hdr"]
        if i > 0:
            s = eig_exact(tr[i - 1]) * r["dt_hdr"] / AS_STAB
            if abs(s / AS_TOL - 1.0) < 1e-6:
                return None
            if s > AS_TOL:
                count = 1 if count < 0 else count + 1
            else:
                count = -1 if count > 0 else count - 1
            if mode == 0 and count > AS_MAXSTIFF:
                mode, dt, up = 1, dt * AS_DTFAC, up + 1
            elif mode == 1 and count < -AS_MAXNONSTIFF:
                mode, dt, down = 0, dt / AS_DTFAC, down + 1
        dt = min(dt, dtmax, tf - r["t"])
        assert r["mode"] == mode, (i, r, mode)
        assert r["dt"] == dt, (i, r, dt)
    return up, down


@pytest.mark.parametrize("alg", ["auto", "auto5p"])
@pytest.mark.parametrize("controller", [0, 1])
def test_autoswitch_switch_to_stiff_pinned_on_linear_decay(oracle, alg, controller):
    """u' = lambda u: Tsit5's estimate |(k7 - k6) / (u1 - g6)| is |lambda|
    (to rounding) and the Rosenbrock one, |J|_inf, is |lambda| exactly; the
    whole method/dt sequence follows from the rule above."""
    decided = 0
    for lam in (-20.0, -150.0, -1000.0, -3.3e3, -2.0e4, -1.0e5):
        r, tr = oracle.solve_trace("exp", [1.0], [lam], (0.0, 3.0), alg=alg, abstol=1e-8,
                                   reltol=1e-6, controller=controller)
        assert r["retcode"] == 1 and len(tr) == r["naccept"] + r["nreject"]
        t0 = tr["mode"] == 0
        np.testing.assert_allclose(tr["eig"][t0], abs(lam), rtol=1e-9)
        assert (tr["eig"][~t0] == abs(lam)).all()
        got = replay_autoswitch(tr, lambda rec: abs(lam), 3.0, 3.0)
        if got is None:
            continue
        decided += 1
        # |lambda| = 20 stays below Tsit5's stability bound at this accuracy
        assert got == ((1, 0) if lam <= -150.0 else (0, 0)) and r["nswitch"] == got[0]
        assert r["nstiff"] == int((tr["accepted"][tr["mode"] == 1]).sum())
    assert decided >= 5


@pytest.mark.parametrize("alg", ["auto", "auto5p"])
def test_autoswitch_switch_back_pinned_on_dropping_stiffness(oracle, alg):
    """A stiff phase that ends: |lambda(t)| drops from ~1e4 to ~1 past
    t = p2.  Estimates: Tsit5's at the step end (stages 6 and 7 both at t +
    dt), the Rosenbrock one |J(t)| at the step start; the rule then fixes one
    switch to the Rosenbrock method and one back to Tsit5 (dt / 2)."""
    import user_model_host as umh
    umh.install(oracle, 1, 3, DROP_RHS, jac=DROP_JAC)
    decided = 0
    for p in DROP_PARAMS:
        p = np.array(p)
        r, tr = oracle.solve_trace("user", [1.0], p, (0.0, 6.0), alg=alg, abstol=1e-8,
                                   reltol=1e-6)
        assert r["retcode"] == 1
        np.testing.assert_allclose(r["u"][0], 1.0 / 37.0, rtol=1e-5)

        def eig_exact(rec):
            tt = rec["t"] + rec["dt"] if rec["mode"] == 0 else rec["t"]
            return abs(lam_drop(p, tt))

        ex = np.array([eig_exact(x) for x in tr])
        np.testing.assert_allclose(tr["eig"], ex, rtol=1e-6)
        got = replay_autoswitch(tr, eig_exact, 6.0, 6.0)
        if got is None:
            continue
        decided += 1
        assert got == (1, 1) and r["nswitch"] == 2
    assert decided >= 2


def test_ensemble_save_layouts(oracle):
    """The oracle's two layouts of an ensemble's saved values hold the same
    numbers ([ns][n][N] and [N][ns][n]); rows past a ragged instance's size
    stay NaN; each trajectory's saves equal its single solve's."""
    from mideq import workloads

    w = workloads.brusselator(6, seed=5)
    tf = np.minimum(w["tf"], 0.5)
    sa = np.array([0.0, 0.1, 0.25, 0.5, 0.7])
    kw = dict(abstol=1e-8, reltol=1e-6, saveat=sa)
    a = oracle.ensemble("brusselator", w["u0"], w["p"], 0.0, tf, nvec=w["nvec"], **kw)
    b = oracle.ensemble("brusselator", w["u0"], w["p"], 0.0, tf, nvec=w["nvec"], save_layout=1,
                        **kw)
    np.testing.assert_array_equal(np.transpose(b["saveat"], (1, 2, 0)), a["saveat"])
    for k in range(6):
        n = int(w["nvec"][k])
        assert np.isnan(a["saveat"][:, n:, k]).all()
        one = oracle.ensemble("brusselator", w["u0"][:n, k:k + 1], w["p"][:, k:k + 1], 0.0,
                              tf[k:k + 1], **kw)
        np.testing.assert_array_equal(a["saveat"][:, :n, k], one["saveat"][:, :, 0])
        assert np.isnan(a["saveat"][sa > tf[k], :n, k]).all()
