"""Uneven Brusselator batch (ragged state sizes 16..128, per-instance
tspans) in the persistent team kernel, against the oracle run with the
same team-size reduction order: identical retcodes and step counts, equal
final states."""
import numpy as np
import pytest

from gpu_util import assert_identical, cpu, gpu, say
from mideq import workloads

pytestmark = pytest.mark.gpu


def batch(ntraj, seed, nmax_grid=64, tf_max=None):
    w = workloads.brusselator(ntraj, seed=seed)
    nvec = w["nvec"].copy()
    if nmax_grid < 64:
        ng = 8 + (nvec // 2 - 8) % (nmax_grid - 7)
        nvec = (2 * ng).astype(np.int32)
        # rebuild initial states for the smaller grids
        u0 = np.zeros_like(w["u0"])
        for k, n in enumerate(nvec):
            x = np.arange(n // 2) / (n // 2)
            u0[0:n:2, k] = 1 + np.sin(2 * np.pi * x)
            u0[1:n:2, k] = 3.0
        w["u0"] = u0
        w["nvec"] = nvec
    if tf_max is not None:
        w["tf"] = np.minimum(w["tf"], tf_max)
    return w


@pytest.mark.parametrize("lanes,precision", [(64, "fp64"), (32, "fp64"), (64, "mixed")])
def test_bruss_tsit5_bitexact(oracle, lanes, precision):
    N = 256
    say(f"brusselator tsit5 {precision} lanes={lanes} N={N}")
    w = batch(N, seed=51, tf_max=5.0)
    tol = dict(abstol=1e-8, reltol=1e-6)
    g = gpu("brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"], lanes=lanes,
            precision=precision, schedule="persistent", **tol)
    c = cpu(oracle, "brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"],
            norm_lanes=lanes, precision=precision, **tol)
    assert (c["retcode"] == 1).all()
    assert g.lanes == lanes
    assert_identical(g, c, nvec=w["nvec"])


@pytest.mark.parametrize("alg,lanes", [("rodas5", 64), ("rodas5", 32), ("auto", 64),
                                       ("rodas5p", 64), ("auto5p", 32)])
def test_bruss_rodas5_and_auto_bitexact(oracle, alg, lanes):
    N = 48
    say(f"brusselator {alg} fp64 lanes={lanes} N={N}")
    w = batch(N, seed=52, nmax_grid=20, tf_max=3.0)
    tol = dict(abstol=1e-8, reltol=1e-6)
    g = gpu("brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"], lanes=lanes,
            alg=alg, schedule="persistent", **tol)
    c = cpu(oracle, "brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"],
            norm_lanes=lanes, alg=alg, **tol)
    assert (c["retcode"] == 1).all()
    assert_identical(g, c, nvec=w["nvec"])
    np.testing.assert_array_equal(g.nstiff, c["nstiff"])


def test_bruss_autotuned_graph_and_order(oracle):
    N = 300
    say(f"brusselator tsit5 autotuned+graph N={N}")
    w = batch(N, seed=53, tf_max=4.0)
    tol = dict(abstol=1e-8, reltol=1e-6)
    g1 = gpu("brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"], lanes=0,
             graph=True, schedule="persistent", **tol)
    assert g1.lanes in (32, 64)
    g2 = gpu("brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"], lanes=0,
             graph=True, schedule="persistent", **tol)
    np.testing.assert_array_equal(g1.u_final, g2.u_final)
    c = cpu(oracle, "brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"],
            norm_lanes=g1.lanes, **tol)
    assert_identical(g1, c, nvec=w["nvec"])
    # natural (unsorted) queue order gives the same per-instance results
    g3 = gpu("brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"], lanes=g1.lanes,
             schedule="lane", **tol)
    assert_identical(g3, c, nvec=w["nvec"])


def test_bruss_invalid_sizes_rejected():
    say("brusselator invalid nvec")
    w = batch(8, seed=54)
    w["nvec"][3] = 15
    with pytest.raises(Exception):
        gpu("brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"], lanes=64)


@pytest.mark.parametrize("alg,lanes,precision", [("tsit5", 64, "fp64"), ("tsit5", 32, "mixed"),
                                                 ("rodas5", 64, "fp64"), ("auto", 32, "fp64")])
def test_bruss_fast_controller_bitexact(oracle, alg, lanes, precision):
    """The team kernels with the division-free controller (EEst^2 = mean r^2,
    log2-domain PI law) against the oracle's restatement."""
    N = 48 if alg != "tsit5" else 200
    say(f"brusselator fast controller {alg} {precision} lanes={lanes} N={N}")
    w = batch(N, seed=55, nmax_grid=20 if alg != "tsit5" else 64, tf_max=3.0)
    tol = dict(abstol=1e-8, reltol=1e-6)
    g = gpu("brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"], lanes=lanes,
            alg=alg, precision=precision, schedule="persistent", controller="fast", **tol)
    c = cpu(oracle, "brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"],
            norm_lanes=lanes, alg=alg, precision=precision, controller=1, **tol)
    assert (c["retcode"] == 1).all()
    assert_identical(g, c, nvec=w["nvec"])
    np.testing.assert_array_equal(g.nstiff, c["nstiff"])


@pytest.mark.parametrize("alg,precision,controller", [("tsit5", "fp64", "reference"),
                                                      ("tsit5", "mixed", "fast"),
                                                      ("auto", "fp64", "reference")])
def test_bruss_narrow_batch_32_lane_shape(oracle, alg, precision, controller):
    """A batch whose layout size is n <= 64 runs the 32-lane teams with one
    (u, v) pair per lane (two instances per wavefront); same results as the
    oracle with the 32-lane reduction order."""
    N = 200 if alg == "tsit5" else 40
    say(f"brusselator n<=64 layout, 32-lane shape, {alg} {precision} {controller} N={N}")
    w = batch(N, seed=60, nmax_grid=32, tf_max=3.0 if alg == "tsit5" else 1.0)
    w["u0"] = np.ascontiguousarray(w["u0"][:64])
    assert w["nvec"].max() <= 64
    tol = dict(abstol=1e-8, reltol=1e-6)
    g = gpu("brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"], lanes=32, alg=alg,
            precision=precision, controller=controller, schedule="persistent", **tol)
    c = cpu(oracle, "brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"], norm_lanes=32,
            alg=alg, precision=precision, controller=1 if controller == "fast" else 0, **tol)
    assert (c["retcode"] == 1).all()
    assert_identical(g, c, nvec=w["nvec"])
    np.testing.assert_array_equal(g.nstiff, c["nstiff"])


def batch_range(ntraj, seed, lo, hi, tf_max):
    """Grid sizes drawn from lo..hi (the first two at hi, so n = 2 hi is
    always present), initial states rebuilt for them."""
    w = workloads.brusselator(ntraj, seed=seed)
    ng = lo + (w["nvec"] // 2 - 8) % (hi - lo + 1)
    ng[:2] = hi
    u0 = np.zeros((2 * hi, ntraj))
    for k, g in enumerate(ng):
        x = np.arange(g) / g
        u0[0:2 * g:2, k] = 1 + np.sin(2 * np.pi * x)
        u0[1:2 * g:2, k] = 3.0
    return dict(u0=u0, p=w["p"], tf=np.minimum(w["tf"], tf_max),
                nvec=(2 * ng).astype(np.int32))


@pytest.mark.parametrize("alg", ["rodas5", "rodas5p", "auto", "auto5p"])
@pytest.mark.parametrize("lanes", [64, 32])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_bruss_rosenbrock_team_full_size_range(oracle, alg, lanes, precision):
    """The Rosenbrock team path over the whole size range of the uneven
    batch: Ngrid 21..64 (n = 42..128, so every row of a 128 x 128 W and every
    lane-strided row slot of both team shapes is factored and solved),
    against the oracle: identical retcodes and step, f-eval and stiff-step
    counts, equal final states."""
    N = 64
    say(f"brusselator {alg} {precision} lanes={lanes} Ngrid 21..64 N={N}")
    w = batch_range(N, seed=70, lo=21, hi=64, tf_max=1.5)
    tol = dict(abstol=1e-8, reltol=1e-6)
    g = gpu("brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"], lanes=lanes,
            alg=alg, precision=precision, schedule="persistent", **tol)
    c = cpu(oracle, "brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"],
            norm_lanes=lanes, alg=alg, precision=precision, **tol)
    assert (c["retcode"] == 1).all()
    if alg.startswith("auto"):
        assert (c["nstiff"] > 0).any() and (c["nstiff"] < c["naccept"]).any()
    assert g.lanes == lanes
    assert_identical(g, c, nvec=w["nvec"])
    np.testing.assert_array_equal(g.nstiff, c["nstiff"])


@pytest.mark.parametrize("alg,lanes,precision", [("rodas5", 64, "fp64"), ("auto", 32, "fp64"),
                                                 ("auto5p", 64, "mixed"),
                                                 ("rodas5p", 32, "mixed")])
def test_bruss_rosenbrock_team_fast_controller_full_size(oracle, alg, lanes, precision):
    """Same size range with the division-free controller."""
    N = 48
    say(f"brusselator fast controller {alg} {precision} lanes={lanes} Ngrid 21..64 N={N}")
    w = batch_range(N, seed=71, lo=21, hi=64, tf_max=1.5)
    tol = dict(abstol=1e-8, reltol=1e-6)
    g = gpu("brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"], lanes=lanes,
            alg=alg, precision=precision, schedule="persistent", controller="fast", **tol)
    c = cpu(oracle, "brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"],
            norm_lanes=lanes, alg=alg, precision=precision, controller=1, **tol)
    assert (c["retcode"] == 1).all()
    assert_identical(g, c, nvec=w["nvec"])
    np.testing.assert_array_equal(g.nstiff, c["nstiff"])


def test_bruss_tsit5_32_lane_fp64_fast_full_size(oracle):
    """The benchmarked instantiation: Tsit5, FP64, division-free controller,
    32-lane teams over the full size range (n up to 128), i.e. the kernel
    built for 4 waves per SIMD."""
    N = 256
    say(f"brusselator tsit5 fp64 fast lanes=32 full range N={N}")
    w = batch(N, seed=72, tf_max=4.0)
    assert w["nvec"].max() == 128
    tol = dict(abstol=1e-8, reltol=1e-6)
    g = gpu("brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"], lanes=32,
            schedule="persistent", controller="fast", **tol)
    c = cpu(oracle, "brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"],
            norm_lanes=32, controller=1, **tol)
    assert (c["retcode"] == 1).all()
    assert g.lanes == 32
    assert_identical(g, c, nvec=w["nvec"])


def soa(us_traj_major):
    """[N, ns, n] -> [ns, n, N]"""
    return np.ascontiguousarray(np.transpose(us_traj_major, (1, 2, 0)))


@pytest.mark.parametrize("alg,lanes,precision,layout", [
    ("tsit5", 64, "fp64", "soa"), ("tsit5", 32, "fp64", "trajectory"),
    ("tsit5", 32, "mixed", "soa"), ("rodas5", 64, "fp64", "trajectory"),
    ("rodas5p", 32, "mixed", "trajectory"), ("auto", 64, "fp64", "soa"),
    ("auto5p", 32, "fp64", "trajectory")])
def test_bruss_saveat_bitexact(oracle, alg, lanes, precision, layout):
    """saveat on the team kernel over the ragged batch (n = 16..128): the
    points each accepted step passes from its dense output (Tsit5's
    interpolant or the Rosenbrock one), u0 at t0, NaN past an instance's tf,
    rows past an instance's size untouched; both layouts, bit-exact against
    the oracle."""
    rosen = alg != "tsit5"
    N = 48 if rosen else 160
    say(f"brusselator saveat {alg} {precision} lanes={lanes} layout={layout} N={N}")
    w = batch_range(N, seed=80, lo=8, hi=64, tf_max=1.5 if rosen else 4.0)
    sa = np.array([0.0, 0.05, 0.3, 0.3, 0.77, 1.0, 1.4999, 2.5, 3.9])
    tol = dict(abstol=1e-8, reltol=1e-6)
    g = gpu("brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"], lanes=lanes, alg=alg,
            precision=precision, schedule="persistent", saveat=sa, save_layout=layout, **tol)
    c = cpu(oracle, "brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"],
            norm_lanes=lanes, alg=alg, precision=precision, saveat=sa, **tol)
    assert (c["retcode"] == 1).all()
    us = soa(g.us) if layout == "trajectory" else g.us
    assert us.shape == c["saveat"].shape
    np.testing.assert_array_equal(us, c["saveat"])
    g.us = None
    assert_identical(g, c, nvec=w["nvec"])
    np.testing.assert_array_equal(g.nstiff, c["nstiff"])
    # u0 at t0 exactly, NaN past tf, rows past n NaN (the buffer's own values)
    n0 = int(w["nvec"][0])
    assert np.array_equal(us[0, :n0, 0], w["u0"][:n0, 0])
    late = sa[None, :] > w["tf"][:, None]
    for k in range(N):
        n = int(w["nvec"][k])
        assert np.isnan(us[late[k], :n, k]).all() and not np.isnan(us[~late[k], :n, k]).any()
        assert np.isnan(us[:, n:, k]).all()


def test_bruss_saveat_f32_storage_and_autotuned(oracle):
    """FP32 storage of the saved values (the FP64 values rounded) under the
    autotuned team shape, with a hipGraph replay giving the same values."""
    N = 200
    say(f"brusselator saveat f32 storage, autotuned, graph N={N}")
    w = batch(N, seed=81, tf_max=3.0)
    sa = np.linspace(0.0, 3.0, 13)
    tol = dict(abstol=1e-8, reltol=1e-6)
    kw = dict(nvec=w["nvec"], lanes=0, schedule="persistent", saveat=sa, save_dtype="float32",
              graph=True, **tol)
    g1 = gpu("brusselator", w["u0"], w["p"], 0.0, w["tf"], **kw)
    g2 = gpu("brusselator", w["u0"], w["p"], 0.0, w["tf"], **kw)
    assert g2.graph
    np.testing.assert_array_equal(g1.us, g2.us)
    c = cpu(oracle, "brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"],
            norm_lanes=g1.lanes, saveat=sa, **tol)
    np.testing.assert_array_equal(g1.us, c["saveat"].astype(np.float32))
    g1.us = None
    assert_identical(g1, c, nvec=w["nvec"])




@pytest.mark.parametrize("alg,lanes,precision,layout", [
    ("tsit5", 64, "fp64", None), ("tsit5", 32, "mixed", "soa"), ("rodas5", 64, "fp64", "soa"),
    ("rodas5", 32, "fp64", None), ("auto", 64, "fp64", "trajectory"),
    ("rodas5p", 32, "mixed", "trajectory")])
def test_bruss_backward_spans_bitexact(oracle, alg, lanes, precision, layout):
    """Instances integrated forward (0 -> tf) and backward (tf -> 0.3 tf, a
    span short enough for the diffusion run in reverse to stay bounded) in
    one ragged batch on the team kernel, plus a zero-length span (Success at
    u0) and a NaN end (InitialFailure at u0); with saveat, backward instances
    fill the points in [tf, t0] from the last one down.  Bit-exact against
    the oracle."""
    N = 48
    say(f"brusselator backward spans {alg} {precision} lanes={lanes} saveat={layout} N={N}")
    w = batch_range(N, seed=83, lo=8, hi=24, tf_max=0.3)
    back = np.arange(N) % 2 == 1
    t0 = np.where(back, w["tf"], 0.0)
    tf = np.where(back, 0.3 * w["tf"], w["tf"])
    tf[4], tf[6] = t0[4], np.nan
    sa = np.linspace(0.0, 0.3, 7) if layout else None
    tol = dict(abstol=1e-8, reltol=1e-6)
    g = gpu("brusselator", w["u0"], w["p"], t0, tf, nvec=w["nvec"], lanes=lanes, alg=alg,
            precision=precision, schedule="persistent", saveat=sa,
            save_layout=layout or "soa", **tol)
    c = cpu(oracle, "brusselator", w["u0"], w["p"], t0, tf, nvec=w["nvec"], norm_lanes=lanes,
            alg=alg, precision=precision, saveat=sa, **tol)
    assert c["retcode"][6] == 6 and (np.delete(c["retcode"], 6) == 1).all()
    assert (c["naccept"][back] > 0).all() and c["naccept"][4] == 0
    if layout:
        us = soa(g.us) if layout == "trajectory" else g.us
        np.testing.assert_array_equal(us, c["saveat"])
        g.us = None
        # backward: points in [0.3 t0, t0] saved, those below NaN
        k = 1
        n = int(w["nvec"][k])
        reached = (sa >= tf[k]) & (sa <= t0[k])
        assert not np.isnan(us[reached, :n, k]).any() and np.isnan(us[~reached, :n, k]).all()
    assert_identical(g, c, nvec=w["nvec"])
    for k in (4, 6):
        n = int(w["nvec"][k])
        np.testing.assert_array_equal(g.u_final[:n, k], w["u0"][:n, k])
    assert (g.t_final[back] == tf[back]).all()


@pytest.mark.parametrize("alg,lanes,precision", [("rodas5", 64, "fp64"), ("auto", 32, "fp64"),
                                                 ("rodas5p", 64, "mixed"), ("rodas5", 32, "fp64")])
def test_bruss_rosenbrock_nonfinite_inputs_bitexact(oracle, alg, lanes, precision):
    """Instances whose W or right-hand side turn non-finite (NaN and infinite
    parameters, a NaN state, states and a parameter large enough to
    overflow, a denormal state) beside ordinary ones: the band LU's outcomes
    (zero pivot first, or a non-finite value meeting a skipped zero of the
    dense LU) and the solves' non-finite checks give the oracle's dense-LU
    retcodes and step counts, bit for bit."""
    N = 12
    say(f"brusselator non-finite inputs {alg} {precision} lanes={lanes} N={N}")
    w = batch_range(N, seed=84, lo=8, hi=14, tf_max=1.0)
    p, u0 = w["p"].copy(), w["u0"].copy()
    p[1, 2] = np.nan
    p[0, 3] = np.inf
    p[2, 4] = np.nan
    u0[3, 5] = np.nan
    u0[:, 6] *= 1e150
    p[1, 7] = 1e200
    u0[0, 8] = 1e-320
    tol = dict(abstol=1e-8, reltol=1e-6)
    g = gpu("brusselator", u0, p, 0.0, w["tf"], nvec=w["nvec"], lanes=lanes, alg=alg,
            precision=precision, schedule="persistent", **tol)
    c = cpu(oracle, "brusselator", u0, p, 0.0, w["tf"], nvec=w["nvec"], norm_lanes=lanes,
            alg=alg, precision=precision, **tol)
    assert (c["retcode"][2:7] == 5).all() and c["retcode"][0] == 1
    assert_identical(g, c, nvec=w["nvec"])
    np.testing.assert_array_equal(g.nstiff, c["nstiff"])
