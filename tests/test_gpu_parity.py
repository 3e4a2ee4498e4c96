"""GPU parity: libmideq.so (HIP, gfx950) against the CPU oracle on the same
seeded ensembles.  FP64 and mixed runs must agree bit for bit: identical
retcodes and accept/reject counts per trajectory, equal final states and
saveat values (the kernels and the oracle follow one arithmetic order)."""
import numpy as np
import pytest

from gpu_util import assert_identical, cpu, gpu, say
from mideq import workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("schedule", ["lane", "persistent"])
def test_lorenz_fp64_saveat_bitexact(oracle, schedule):
    N = 3000
    say(f"lorenz fp64 tsit5 N={N} saveat=100 schedule={schedule}")
    w = workloads.lorenz(N, seed=11)
    sa = workloads.saveat_uniform(0.0, 10.0, 100)
    g = gpu("lorenz", w["u0"], w["p"], 0.0, 10.0, saveat=sa, abstol=1e-8, reltol=1e-8,
            schedule=schedule)
    c = cpu(oracle, "lorenz", w["u0"], w["p"], 0.0, 10.0, saveat=sa, abstol=1e-8, reltol=1e-8)
    assert (c["retcode"] == 1).all()
    assert_identical(g, c)
    # endpoints of the saveat grid: u0 copied exactly, last point = final state
    np.testing.assert_array_equal(g.us[0], w["u0"])
    np.testing.assert_array_equal(g.us[-1], g.u_final)


def test_lorenz_fp64_float32_output(oracle):
    N = 1000
    say(f"lorenz fp64 f32-saveat N={N}")
    w = workloads.lorenz(N, seed=12)
    sa = workloads.saveat_uniform(0.0, 10.0, 100)
    g = gpu("lorenz", w["u0"], w["p"], 0.0, 10.0, saveat=sa, abstol=1e-8, reltol=1e-8,
            save_dtype="float32")
    c = cpu(oracle, "lorenz", w["u0"], w["p"], 0.0, 10.0, saveat=sa, abstol=1e-8, reltol=1e-8)
    assert g.us.dtype == np.float32
    assert_identical(g, c, f32=True)


@pytest.mark.parametrize("schedule", ["lane", "persistent"])
def test_lorenz_mixed_bitexact(oracle, schedule):
    N = 3000
    say(f"lorenz mixed tsit5 N={N} schedule={schedule}")
    w = workloads.lorenz(N, seed=13)
    sa = workloads.saveat_uniform(0.0, 10.0, 100)
    g = gpu("lorenz", w["u0"], w["p"], 0.0, 10.0, saveat=sa, abstol=1e-6, reltol=1e-6,
            precision="mixed", schedule=schedule)
    c = cpu(oracle, "lorenz", w["u0"], w["p"], 0.0, 10.0, saveat=sa, abstol=1e-6, reltol=1e-6,
            precision="mixed")
    assert_identical(g, c)


@pytest.mark.parametrize("alg", ["rodas5", "rodas5p"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_robertson_rodas5_bitexact(oracle, precision, alg):
    N = 2048
    say(f"robertson {alg} {precision} N={N}")
    w = workloads.robertson(N, seed=21)
    g = gpu("robertson", w["u0"], w["p"], 0.0, 1e5, alg=alg, precision=precision,
            abstol=1e-8, reltol=1e-6)
    c = cpu(oracle, "robertson", w["u0"], w["p"], 0.0, 1e5, alg=alg, precision=precision,
            abstol=1e-8, reltol=1e-6)
    assert (c["retcode"] == 1).all()
    assert_identical(g, c)
    if precision == "fp64":
        assert np.abs(g.u_final.sum(axis=0) - 1.0).max() < 1e-12


@pytest.mark.parametrize("model,alg", [("vdp", "auto"), ("vdp", "rodas5"), ("robertson", "auto"),
                                       ("lorenz", "auto"), ("lorenz", "rodas5"),
                                       ("vdp", "auto5p"), ("vdp", "rodas5p"),
                                       ("robertson", "auto5p"), ("lorenz", "rodas5p")])
def test_auto_and_rodas_models_bitexact(oracle, model, alg):
    N = 256
    say(f"{model} {alg} fp64 N={N}")
    if model == "vdp":
        w = workloads.vdp(N, seed=31)
        w["p"] = np.minimum(w["p"], 300.0)  # keep the oracle's share short
        tf = 2.0 * w["p"][0]
        tol = dict(abstol=1e-8, reltol=1e-6)
    elif model == "robertson":
        w = workloads.robertson(N, seed=32)
        tf = 1e5
        tol = dict(abstol=1e-8, reltol=1e-6)
    else:
        w = workloads.lorenz(N, seed=33)
        tf = 10.0
        tol = dict(abstol=1e-7, reltol=1e-7)
    g = gpu(model, w["u0"], w["p"], 0.0, tf, alg=alg, **tol)
    c = cpu(oracle, model, w["u0"], w["p"], 0.0, tf, alg=alg, **tol)
    assert_identical(g, c)
    np.testing.assert_array_equal(g.nstiff, c["nstiff"])


def test_linear_and_exp_models(oracle):
    rng = np.random.default_rng(5)
    N = 512
    say(f"linear3 tsit5/rodas5 + exp N={N}")
    M = rng.uniform(-1, 0.5, (9, N))
    u0 = rng.uniform(-1, 1, (3, N))
    for alg in ("tsit5", "rodas5"):
        g = gpu("linear", u0, M, 0.0, 3.0, alg=alg, abstol=1e-9, reltol=1e-9)
        c = cpu(oracle, "linear", u0, M, 0.0, 3.0, alg=alg, abstol=1e-9, reltol=1e-9)
        assert_identical(g, c)
    lam = rng.uniform(-2, 2, (1, N))
    u0 = rng.uniform(0.1, 1, (1, N))
    g = gpu("exp", u0, lam, 0.0, 1.0, abstol=1e-8, reltol=1e-8)
    c = cpu(oracle, "exp", u0, lam, 0.0, 1.0, abstol=1e-8, reltol=1e-8)
    assert_identical(g, c)
    np.testing.assert_allclose(g.u_final[0], u0[0] * np.exp(lam[0]), rtol=1e-7)


def test_retcodes_forced_from_host(oracle):
    N = 64
    say("retcodes maxiters/dtmin/nan")
    w = workloads.lorenz(N, seed=41)
    g = gpu("lorenz", w["u0"], w["p"], 0.0, 10.0, maxiters=5)
    assert (g.retcode == 2).all() and ((g.naccept + g.nreject) == 5).all()
    g = gpu("lorenz", w["u0"], w["p"], 0.0, 10.0, dtmin=0.5)
    assert (g.retcode == 3).all()
    p = w["p"].copy()
    p[0, 3] = np.nan
    g = gpu("lorenz", w["u0"], p, 0.0, 1.0)
    c = cpu(oracle, "lorenz", w["u0"], p, 0.0, 1.0)
    assert g.retcode[3] in (4, 5) and (np.delete(g.retcode, 3) == 1).all()
    np.testing.assert_array_equal(g.retcode, c["retcode"])
    # per-trajectory spans: zero length -> Success at u0; reversed -> integrated
    # backward to tf; a NaN end -> InitialFailure at u0
    tf = np.full(N, 1.0)
    tf[5], tf[9], tf[11] = 0.0, -1.0, np.nan
    g = gpu("lorenz", w["u0"], w["p"], 0.0, tf)
    c = cpu(oracle, "lorenz", w["u0"], w["p"], 0.0, tf)
    assert g.retcode[5] == 1 and g.retcode[9] == 1 and g.retcode[11] == 6
    assert g.t_final[9] == -1.0 and g.naccept[9] > 0
    np.testing.assert_array_equal(g.u_final[:, [5, 11]], w["u0"][:, [5, 11]])
    assert_identical(g, c)


@pytest.mark.parametrize("alg,precision,controller,schedule", [
    ("tsit5", "fp64", "reference", "lane"), ("tsit5", "mixed", "fast", "persistent"),
    ("rodas5", "fp64", "fast", "lane"), ("auto", "fp64", "reference", "lane"),
    ("rodas5p", "mixed", "reference", "lane")])
def test_backward_and_forward_spans_with_saveat(oracle, alg, precision, controller, schedule):
    """Trajectories integrated forward and backward in one ensemble against
    one ascending saveat: backward ones fill the points in (tf, t0] from the
    last one down; bit-exact against the oracle."""
    N = 1000
    say(f"backward spans {alg} {precision} {controller} {schedule} N={N}")
    rng = np.random.default_rng(45)
    # decay forward, growth backward: both directions stay well posed (the
    # Lorenz flow run backward from the attractor's neighbourhood is not)
    model, w = "exp", dict(u0=np.ones((1, N)), p=-rng.uniform(0.5, 2.0, (1, N)))
    back = rng.random(N) < 0.5
    t0 = np.where(back, 1.0, 0.0)
    tf = np.where(back, 0.0, 1.0) + rng.uniform(-0.3, 0.3, N)
    sa = np.linspace(-0.5, 1.5, 11)
    kw = dict(alg=alg, precision=precision, controller=controller, schedule=schedule, saveat=sa,
              abstol=1e-8, reltol=1e-6)
    g = gpu(model, w["u0"], w["p"], t0, tf, **kw)
    c = cpu(oracle, model, w["u0"], w["p"], t0, tf, alg=alg, precision=precision,
            controller=1 if controller == "fast" else 0, saveat=sa, abstol=1e-8, reltol=1e-6)
    assert (c["retcode"] == 1).all()
    assert_identical(g, c, f32=False)
    np.testing.assert_array_equal(g.t_final, tf)
    # saved points strictly outside [min(t0, tf), max(t0, tf)] stay NaN
    lo, hi = np.minimum(t0, tf), np.maximum(t0, tf)
    outside = (sa[:, None] < lo[None, :]) | (sa[:, None] > hi[None, :])
    assert np.isnan(g.us[:, 0, :][outside]).all() and not np.isnan(g.us[:, 0, :][~outside]).any()
    inside = g.us[:, 0, :][~outside]
    exact = np.exp(w["p"][0][None, :] * (sa[:, None] - t0[None, :]))[~outside]
    np.testing.assert_allclose(inside, exact, rtol=1e-4)


@pytest.mark.parametrize("alg,precision,controller", [
    ("tsit5", "fp64", "reference"), ("tsit5", "mixed", "fast"), ("rodas5", "fp64", "fast"),
    ("auto", "fp64", "reference")])
def test_backward_span_ending_near_zero_snaps_to_tf(oracle, alg, precision, controller):
    """Backward spans whose end is tiny next to their start, e.g. (1, 1e-3)
    and (10, 1e-6): the last step's end snaps to tf within 100 eps(max(t,
    tf)) = 100 eps(t), t the step's start (a window of 100 eps(tf) would miss
    and overshoot).  t_final == tf and results bit-exact vs the oracle."""
    N = 512
    say(f"backward spans ending near 0 {alg} {precision} {controller} N={N}")
    rng = np.random.default_rng(46)
    u0 = np.ones((1, N))
    p = rng.uniform(0.5, 2.0, (1, N))  # u' = p u run backward: decay
    t0 = rng.choice([1.0, 2.0, 10.0], N) * rng.uniform(0.9, 1.1, N)
    tf = rng.choice([1e-3, 1e-6, 1e-9, 0.0, -1e-7], N) * rng.uniform(0.5, 2.0, N)
    kw = dict(abstol=1e-8, reltol=1e-6)
    g = gpu("exp", u0, p, t0, tf, alg=alg, precision=precision, controller=controller, **kw)
    c = cpu(oracle, "exp", u0, p, t0, tf, alg=alg, precision=precision,
            controller=1 if controller == "fast" else 0, **kw)
    assert (c["retcode"] == 1).all()
    np.testing.assert_array_equal(c["t"], tf)
    assert_identical(g, c)
    np.testing.assert_array_equal(g.t_final, tf)


def test_per_trajectory_tspans_and_fixed_dt(oracle):
    N = 500
    say(f"per-trajectory tspan N={N}")
    w = workloads.lorenz(N, seed=42)
    rng = np.random.default_rng(1)
    t0 = rng.uniform(0, 1, N)
    tf = t0 + rng.uniform(0.5, 3, N)
    g = gpu("lorenz", w["u0"], w["p"], t0, tf, abstol=1e-9, reltol=1e-9)
    c = cpu(oracle, "lorenz", w["u0"], w["p"], t0, tf, abstol=1e-9, reltol=1e-9)
    assert_identical(g, c)
    np.testing.assert_array_equal(g.t_final, tf)
    g = gpu("lorenz", w["u0"], w["p"], 0.0, 1.0, dt=1e-3, abstol=1e-9, reltol=1e-9)
    c = cpu(oracle, "lorenz", w["u0"], w["p"], 0.0, 1.0, dt=1e-3, abstol=1e-9, reltol=1e-9)
    assert_identical(g, c)


@pytest.mark.parametrize("model,alg,precision", [("lorenz", "tsit5", "fp64"), ("lorenz", "tsit5", "mixed"),
                                                 ("robertson", "rodas5", "fp64"), ("vdp", "auto", "fp64"),
                                                 ("robertson", "auto", "mixed"),
                                                 ("robertson", "rodas5p", "fp64"),
                                                 ("vdp", "auto5p", "fp64")])
def test_fast_controller_bitexact(oracle, model, alg, precision):
    """The division-free controller (EEst^2 in log2) against the oracle's
    restatement of the same form."""
    N = 1024 if model != "vdp" else 256
    say(f"fast controller {model} {alg} {precision} N={N}")
    if model == "lorenz":
        w = workloads.lorenz(N, seed=61)
        tf, tol, sa = 10.0, dict(abstol=1e-8, reltol=1e-8), workloads.saveat_uniform(0.0, 10.0, 100)
    elif model == "robertson":
        w = workloads.robertson(N, seed=62)
        tf, tol, sa = 1e5, dict(abstol=1e-8, reltol=1e-6), None
    else:
        w = workloads.vdp(N, seed=63)
        w["p"] = np.minimum(w["p"], 300.0)
        tf, tol, sa = 2.0 * w["p"][0], dict(abstol=1e-8, reltol=1e-6), None
    g = gpu(model, w["u0"], w["p"], 0.0, tf, alg=alg, precision=precision, saveat=sa,
            controller="fast", **tol)
    c = cpu(oracle, model, w["u0"], w["p"], 0.0, tf, alg=alg, precision=precision, saveat=sa,
            controller=1, **tol)
    assert (c["retcode"] == 1).all()
    assert_identical(g, c)


@pytest.mark.parametrize("controller", ["reference", "fast"])
def test_cfg1_lorenz_10000_endpoint_bitexact(oracle, controller):
    """The reference's own CPU-sized case: 10,000 seeded Lorenz trajectories
    (sigma, rho, beta +-10 %), u0 = [1, 0, 0], tspan (0, 10), Tsit5,
    abstol = reltol = 1e-8, endpoint only -- identical to the oracle's
    EnsembleThreads restatement trajectory by trajectory."""
    N = 10_000
    say(f"cfg1 lorenz N={N} endpoint controller={controller}")
    w = workloads.lorenz(N, seed=1234)
    g = gpu("lorenz", w["u0"], w["p"], 0.0, 10.0, abstol=1e-8, reltol=1e-8,
            controller=controller)
    c = cpu(oracle, "lorenz", w["u0"], w["p"], 0.0, 10.0, abstol=1e-8, reltol=1e-8,
            controller=1 if controller == "fast" else 0)
    assert (c["retcode"] == 1).all()
    assert_identical(g, c)
    assert (g.t_final == 10.0).all()


@pytest.mark.parametrize("ns", [2048, 2500])
def test_many_saveat_points_bitexact(oracle, ns):
    """Save times staged in LDS (ns <= 2048) and read from global memory
    (ns > 2048) give the same values as the oracle."""
    N = 300
    say(f"lorenz saveat ns={ns} N={N}")
    w = workloads.lorenz(N, seed=81)
    sa = workloads.saveat_uniform(0.0, 10.0, ns)
    g = gpu("lorenz", w["u0"], w["p"], 0.0, 10.0, saveat=sa, abstol=1e-8, reltol=1e-8,
            controller="fast")
    c = cpu(oracle, "lorenz", w["u0"], w["p"], 0.0, 10.0, saveat=sa, abstol=1e-8, reltol=1e-8,
            controller=1)
    assert_identical(g, c)


def test_saveat_outside_span_and_duplicates(oracle):
    """Points before t0 and after tf are NaN, a point at t0 is u0 exactly,
    duplicated points repeat the same value."""
    N = 200
    say(f"lorenz saveat edges N={N}")
    w = workloads.lorenz(N, seed=82)
    sa = np.array([-1.0, 0.0, 0.5, 0.5, 1.25, 2.0, 3.5])
    g = gpu("lorenz", w["u0"], w["p"], 0.0, 2.0, saveat=sa, abstol=1e-8, reltol=1e-8)
    c = cpu(oracle, "lorenz", w["u0"], w["p"], 0.0, 2.0, saveat=sa, abstol=1e-8, reltol=1e-8)
    assert_identical(g, c)
    assert np.isnan(g.us[0]).all() and np.isnan(g.us[6]).all()
    np.testing.assert_array_equal(g.us[1], w["u0"])
    np.testing.assert_array_equal(g.us[2], g.us[3])
    np.testing.assert_array_equal(g.us[5], g.u_final)


@pytest.mark.parametrize("model,alg,precision,schedule,save_dtype", [
    ("lorenz", "tsit5", "fp64", "lane", "float64"), ("lorenz", "tsit5", "mixed", "persistent",
                                                     "float32"),
    ("robertson", "rodas5", "fp64", "lane", "float64"), ("vdp", "auto", "fp64", "lane", "float64"),
    ("exp", "rodas5p", "fp64", "lane", "float32")])
def test_trajectory_major_saves_bitexact(oracle, model, alg, precision, schedule, save_dtype):
    """save_layout "trajectory" ([N, ns, n]: one trajectory's saved states
    contiguous) on the one-lane kernels: the same values as the [ns, n, N]
    layout and as the oracle in its trajectory-major layout, NaN at points
    outside a span; the exp model runs half its trajectories backward."""
    N = 700
    say(f"trajectory-major saves {model} {alg} {precision} {schedule} {save_dtype} N={N}")
    rng = np.random.default_rng(46)
    if model == "exp":
        w = dict(u0=np.ones((1, N)), p=-rng.uniform(0.5, 2.0, (1, N)))
        back = rng.random(N) < 0.5
        t0, tf = np.where(back, 1.0, 0.0), np.where(back, 0.0, 1.0) + rng.uniform(-0.3, 0.3, N)
        sa = np.linspace(-0.5, 1.5, 11)
    else:
        w = dict(lorenz=workloads.lorenz, robertson=workloads.robertson,
                 vdp=workloads.vdp)[model](N, seed=47)
        t0 = 0.0
        tf = dict(lorenz=2.0, robertson=10.0, vdp=3.0)[model]
        sa = np.linspace(0.0, tf, 17)[1:]
    tol = dict(abstol=1e-8, reltol=1e-6)
    kw = dict(alg=alg, precision=precision, schedule=schedule, saveat=sa, save_dtype=save_dtype,
              **tol)
    g = gpu(model, w["u0"], w["p"], t0, tf, save_layout="trajectory", **kw)
    s = gpu(model, w["u0"], w["p"], t0, tf, **kw)
    n = w["u0"].shape[0]
    assert g.us.shape == (N, len(sa), n)
    np.testing.assert_array_equal(np.transpose(g.us, (1, 2, 0)), s.us)
    c = cpu(oracle, model, w["u0"], w["p"], t0, tf, alg=alg, precision=precision, saveat=sa,
            save_layout=1, **tol)
    ref = c["saveat"].astype(np.float32) if save_dtype == "float32" else c["saveat"]
    np.testing.assert_array_equal(g.us, ref)
    g.us = None
    assert_identical(g, c)
