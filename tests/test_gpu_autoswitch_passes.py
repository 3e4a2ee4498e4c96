"""AutoTsit5 in two passes (mideq_options.auto_passes): a Tsit5 pass solves
the trajectories that never switch and hands the rest to the full
AutoSwitch kernel, which solves them from t0.  Each trajectory takes the
same steps either way, so every output must be bit-equal to the single
AutoSwitch pass -- over batches where none, some or all trajectories
switch, with stiff-start hints (handed over before the first step), caller
launch orders, both schedules, both precisions, both controllers, ragged
batch sizes and captured-graph replays."""
import numpy as np
import pytest

from gpu_util import assert_identical, cpu, gpu, say
from mideq import workloads

pytestmark = pytest.mark.gpu

KEYS = ("retcode", "naccept", "nreject", "nf", "nstiff", "u_final", "t_final", "dt_first")


def _same(a, b):
    for k in KEYS:
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=k)


def _vdp_mixed(N, seed):
    # mu from 1 (never switches) to 300 (switches early)
    w = workloads.vdp(N, seed=seed)
    w["p"] = np.minimum(w["p"], 300.0)
    return w, 2.0 * w["p"][0]


@pytest.mark.parametrize("alg", ["auto", "auto5p"])
@pytest.mark.parametrize("schedule", ["lane", "persistent"])
@pytest.mark.parametrize("precision,controller", [("fp64", "reference"), ("fp64", "fast"),
                                                  ("mixed", "fast")])
def test_vdp_two_pass_equals_single_pass(alg, schedule, precision, controller):
    N = 4099  # a partial last wavefront
    say(f"vdp {alg} {schedule} {precision} {controller} two-pass vs one N={N}")
    w, tf = _vdp_mixed(N, 81)
    kw = dict(alg=alg, schedule=schedule, precision=precision, controller=controller,
              abstol=1e-8, reltol=1e-6)
    one = gpu("vdp", w["u0"], w["p"], 0.0, tf, auto_passes=1, **kw)
    two = gpu("vdp", w["u0"], w["p"], 0.0, tf, auto_passes=2, **kw)
    nsw = int((one.nstiff > 0).sum())
    assert 0 < nsw < N, nsw  # both passes do work
    _same(one, two)


@pytest.mark.parametrize("model", ["lorenz", "robertson"])
def test_none_or_all_switch(model):
    N = 3000
    say(f"{model} auto two-pass vs one N={N}")
    w = getattr(workloads, model)(N, seed=82)
    tf = 10.0 if model == "lorenz" else 1.0e5
    kw = dict(alg="auto", abstol=1e-8, reltol=1e-6)
    one = gpu(model, w["u0"], w["p"], 0.0, tf, auto_passes=1, **kw)
    two = gpu(model, w["u0"], w["p"], 0.0, tf, auto_passes=2, **kw)
    _same(one, two)
    if model == "robertson":
        assert (one.nstiff > 0).all()


def test_hints_and_order_two_pass(oracle):
    """Stiff-start hints are handed over before any step; a caller order
    applies to the first pass.  Checked against the oracle too."""
    N = 1500
    say(f"vdp auto hints + order two-pass N={N}")
    w, tf = _vdp_mixed(N, 83)
    rng = np.random.default_rng(4)
    hint = (rng.uniform(size=N) > 0.7).astype(np.int8)
    dt0 = 10.0 ** rng.uniform(-6, -2, N)
    kw = dict(abstol=1e-8, reltol=1e-6, alg_hint=hint, dt0=dt0)
    order = rng.permutation(N)
    one = gpu("vdp", w["u0"], w["p"], 0.0, tf, alg="auto", auto_passes=1, **kw)
    two = gpu("vdp", w["u0"], w["p"], 0.0, tf, alg="auto", order=order, auto_passes=2, **kw)
    _same(one, two)
    c = cpu(oracle, "vdp", w["u0"], w["p"], 0.0, tf, alg="auto", **kw)
    assert_identical(two, c)
    np.testing.assert_array_equal(two.nstiff, c["nstiff"])


@pytest.mark.parametrize("N", [1, 63, 65])
def test_small_batches(N):
    say(f"vdp auto two-pass small N={N}")
    w, tf = _vdp_mixed(N, 84)
    kw = dict(alg="auto", abstol=1e-8, reltol=1e-6)
    _same(gpu("vdp", w["u0"], w["p"], 0.0, tf, auto_passes=1, **kw),
          gpu("vdp", w["u0"], w["p"], 0.0, tf, auto_passes=2, **kw))


def test_graph_replays_two_pass():
    """The hand-off counter is cleared by a kernel inside the captured
    sequence, so every replay starts from an empty list."""
    N = 2000
    say(f"vdp auto two-pass graph replays N={N}")
    w, tf = _vdp_mixed(N, 85)
    kw = dict(alg="auto", abstol=1e-8, reltol=1e-6)
    ref = gpu("vdp", w["u0"], w["p"], 0.0, tf, **kw)  # default: one pass
    for rep in range(3):
        g = gpu("vdp", w["u0"], w["p"], 0.0, tf, graph=True, auto_passes=2, **kw)
        _same(ref, g)


@pytest.mark.parametrize("alg", ["auto", "auto5p"])
def test_trust_pass(alg):
    """auto_passes = 3 (trust the hints): a trajectory hinted non-stiff that
    finishes within its cap has plain Tsit5's result; every other one has
    the single AutoSwitch pass's result for the same hints.  Some stiff
    trajectories are hinted non-stiff on purpose, so the cap hands them on."""
    N = 4099
    say(f"vdp {alg} trust pass N={N}")
    w, tf = _vdp_mixed(N, 86)
    kw = dict(abstol=1e-8, reltol=1e-6, maxiters=1_000_000)
    plain = gpu("vdp", w["u0"], w["p"], 0.0, tf, alg=alg, **kw)
    tsit = gpu("vdp", w["u0"], w["p"], 0.0, tf, alg="tsit5", **kw)
    hint = (plain.nstiff > 0).astype(np.int8)
    rng = np.random.default_rng(3)
    wrong = rng.choice(np.where(hint == 1)[0], 200, replace=False)
    hint[wrong] = 0
    cap = (3 * (plain.naccept + plain.nreject) + 16).astype(np.int32)
    auto_h = gpu("vdp", w["u0"], w["p"], 0.0, tf, alg=alg, alg_hint=hint, **kw)
    trust = gpu("vdp", w["u0"], w["p"], 0.0, tf, alg=alg, alg_hint=hint, iter_cap=cap,
                auto_passes=3, **kw)
    first = (hint == 0) & (tsit.naccept + tsit.nreject <= cap)
    assert first.sum() > 100 and (~first).sum() > 100
    capped = (hint == 0) & ~first
    assert capped.sum() > 0  # the cap caught mispredicted trajectories
    for k in KEYS:
        t, a, b = getattr(trust, k), getattr(tsit, k), getattr(auto_h, k)
        sel = (slice(None), first) if t.ndim == 2 else first
        np.testing.assert_array_equal(t[sel], a[sel], err_msg=k)
        sel = (slice(None), ~first) if t.ndim == 2 else ~first
        np.testing.assert_array_equal(t[sel], b[sel], err_msg=k)
