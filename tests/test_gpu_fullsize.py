"""Parity at the benchmark sizes: each configuration runs at its full size on
the GPU (one launch), and a strided sample of its trajectories is compared
with the oracle bit for bit -- the sample's inputs are the same columns, so
any indexing or offset error at full size shows up.  Size-independent
properties are checked over every trajectory (all succeed, t_final == tf,
Robertson's sum y = 1, saveat endpoints)."""
import numpy as np
import pytest

from gpu_util import assert_identical, cpu, gpu, say
from mideq import workloads

pytestmark = pytest.mark.gpu


def _sample(N, k):
    idx = np.unique(np.linspace(0, N - 1, k).astype(np.int64))
    return idx


class _Sub:
    """The sampled columns of a GPU result, shaped like a result."""

    def __init__(self, g, idx):
        for name in ("retcode", "naccept", "nreject", "nf", "t_final", "dt_first"):
            setattr(self, name, getattr(g, name)[idx])
        self.u_final = g.u_final[:, idx]
        self.us = None if g.us is None else g.us[:, :, idx]


def test_lorenz_1M_saveat100_tuned_sampled_bitexact(oracle):
    N = 1 << 20
    say(f"full size lorenz N={N} saveat 100, fast controller")
    w = workloads.lorenz(N, seed=1234)
    sa = workloads.saveat_uniform(0.0, 10.0, 100)
    g = gpu("lorenz", w["u0"], w["p"], 0.0, 10.0, saveat=sa, abstol=1e-8, reltol=1e-8,
            controller="fast")
    assert (g.retcode == 1).all() and (g.t_final == 10.0).all()
    np.testing.assert_array_equal(g.us[0], w["u0"])
    np.testing.assert_array_equal(g.us[-1], g.u_final)
    idx = _sample(N, 1000)
    c = cpu(oracle, "lorenz", w["u0"][:, idx], w["p"][:, idx], 0.0, 10.0, saveat=sa,
            abstol=1e-8, reltol=1e-8, controller=1)
    assert_identical(_Sub(g, idx), c)


def test_robertson_262144_sampled_bitexact_and_invariant(oracle):
    N = 262144
    say(f"full size robertson N={N} rodas5")
    w = workloads.robertson(N, seed=1234)
    g = gpu("robertson", w["u0"], w["p"], 0.0, 1e5, alg="rodas5", abstol=1e-8, reltol=1e-6)
    assert (g.retcode == 1).all() and (g.t_final == 1e5).all()
    assert np.abs(g.u_final.sum(axis=0) - 1.0).max() < 1e-12
    idx = _sample(N, 2000)
    c = cpu(oracle, "robertson", w["u0"][:, idx], w["p"][:, idx], 0.0, 1e5, alg="rodas5",
            abstol=1e-8, reltol=1e-6)
    assert_identical(_Sub(g, idx), c)


def test_brusselator_131072_sampled_bitexact(oracle):
    N = 131072
    say(f"full size brusselator N={N} tsit5 lanes=64 fast controller")
    w = workloads.brusselator(N, seed=1234)
    g = gpu("brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"], lanes=64,
            schedule="persistent", controller="fast", abstol=1e-8, reltol=1e-6)
    assert (g.retcode == 1).all()
    np.testing.assert_array_equal(g.t_final, w["tf"])
    idx = _sample(N, 64)
    c = cpu(oracle, "brusselator", w["u0"][:, idx], w["p"][:, idx], 0.0, w["tf"][idx],
            nvec=w["nvec"][idx], norm_lanes=64, controller=1, abstol=1e-8, reltol=1e-6)
    assert_identical(_Sub(g, idx), c, nvec=w["nvec"][idx])
    # rows past each instance's size are left as the host supplied them (zero)
    mask = np.arange(g.u_final.shape[0])[:, None] >= w["nvec"][None, :]
    assert (g.u_final[mask] == 0.0).all()


def test_vdp_mix_third_learned_inputs_sampled_bitexact(oracle):
    """The Van der Pol third of the 1e6 mix (333,333 trajectories, mu up to
    1e4, tspan (0, 2 mu)) solved the way the learned form runs it --
    AutoTsit5 with per-trajectory dt0 and method hints, a launch order and
    the fast controller -- with a strided sample checked against the oracle
    given the same dt0 and hints."""
    N = 333_333
    say(f"full size vdp N={N} auto, dt0 + hint + order, fast controller")
    w = workloads.vdp(N, seed=1234)
    tf = np.asarray(w["tf"], dtype=np.float64)
    rng = np.random.default_rng(10)
    mu = w["p"][0]
    dt0 = 10.0 ** (rng.uniform(-5, -3, N) - 0.5 * np.log10(mu))
    hint = (mu > 100.0).astype(np.int8)
    order = np.argsort(-np.log(mu), kind="stable")
    kw = dict(abstol=1e-8, reltol=1e-6, dt0=dt0, alg_hint=hint, maxiters=1_000_000)
    g = gpu("vdp", w["u0"], w["p"], 0.0, tf, alg="auto", controller="fast", order=order, **kw)
    assert (g.retcode == 1).all()
    np.testing.assert_array_equal(g.t_final, tf)
    idx = _sample(N, 300)
    kws = dict(abstol=1e-8, reltol=1e-6, dt0=dt0[idx], alg_hint=hint[idx], maxiters=1_000_000)
    c = cpu(oracle, "vdp", w["u0"][:, idx], w["p"][:, idx], 0.0, tf[idx], alg="auto",
            controller=1, **kws)
    assert_identical(_Sub(g, idx), c)

