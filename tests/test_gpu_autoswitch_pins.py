"""AutoTsit5 switching on problems whose stiffness is known in closed form,
GPU against the oracle.  The oracle's switching sequence on the same
problems is pinned by replaying the switching rule with the exact estimate
|lambda| (tests/test_oracle_solver.py); here the kernels must reproduce the
oracle bit for bit: retcodes, accepted / rejected / stiff step counts, f
evaluations and final states."""
import numpy as np
import pytest

import user_model_host as umh
from gpu_util import assert_identical, cpu, gpu, say
from mideq import UserModel
from test_oracle_solver import DROP_JAC, DROP_PARAMS, DROP_RHS
from test_gpu_user_model import run

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("alg,precision,controller", [
    ("auto", "fp64", "reference"), ("auto", "fp64", "fast"), ("auto5p", "fp64", "reference"),
    ("auto", "mixed", "fast")])
def test_linear_decay_switch_to_stiff(oracle, alg, precision, controller):
    """u' = lambda u, lambda from -20 to -1e5: trajectories with |lambda|
    past Tsit5's stability bound at this accuracy switch once."""
    N = 512
    say(f"autoswitch pin u'=lambda u {alg} {precision} {controller} N={N}")
    lam = -np.logspace(np.log10(20.0), 5.0, N)
    u0 = np.ones((1, N))
    p = lam[None, :].copy()
    kw = dict(abstol=1e-8, reltol=1e-6)
    g = gpu("exp", u0, p, 0.0, 3.0, alg=alg, precision=precision, controller=controller, **kw)
    c = cpu(oracle, "exp", u0, p, 0.0, 3.0, alg=alg, precision=precision,
            controller=1 if controller == "fast" else 0, **kw)
    assert (c["retcode"] == 1).all()
    assert (c["nstiff"] > 0).mean() > 0.5 and (c["nstiff"] == 0).any()
    assert_identical(g, c)
    np.testing.assert_array_equal(g.nstiff, c["nstiff"])


@pytest.mark.parametrize("alg", ["auto", "auto5p"])
def test_dropping_stiffness_switch_back(oracle, alg):
    """u' = lambda(t) (u - g) + g', |lambda| dropping from ~1e4 to ~1: one
    switch to the Rosenbrock method and one back to Tsit5 per trajectory
    (user model compiled at run time; lambda rational in t, so host and
    device evaluate f and J identically)."""
    N = 192
    say(f"autoswitch pin dropping stiffness {alg} N={N}")
    rng = np.random.default_rng(80)
    base = np.array(DROP_PARAMS, dtype=np.float64)
    p = base[rng.integers(0, len(base), N)].T * rng.uniform(0.8, 1.25, (3, N))
    u0 = np.ones((1, N))
    umh.install(oracle, 1, 3, DROP_RHS, jac=DROP_JAC)
    m = UserModel(1, 3, DROP_RHS, jac=DROP_JAC, name="drop_stiffness")
    kw = dict(abstol=1e-8, reltol=1e-6)
    g = run(m, u0, p, 0.0, 6.0, alg=alg, **kw)
    c = cpu(oracle, "user", u0, p, 0.0, 6.0, alg=alg, **kw)
    assert (c["retcode"] == 1).all()
    # stiff phase then Tsit5 again: some accepted steps in each method
    assert ((c["nstiff"] > 0) & (c["nstiff"] < c["naccept"])).mean() > 0.9
    assert_identical(g, c)
    np.testing.assert_array_equal(g.nstiff, c["nstiff"])
