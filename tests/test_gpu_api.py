"""The SciMLBase-style surface on the GPU: EnsembleProblem with prob_func,
output_func (rerun) and reduction, and the reference's known answers
(test/core.jl)."""
import math

import numpy as np
import pytest

import mideq
from gpu_util import say
from mideq import (EnsembleMI355X, EnsembleProblem, ODEProblem, ReturnCode, Rodas5, Tsit5,
                   remake, solve)

pytestmark = pytest.mark.gpu


def test_core_basic_solve_on_gpu():
    say("core.jl basic ODE solve")
    prob = ODEProblem(mideq.exp_growth, 0.5, (0.0, 1.0), [1.01])
    sol = solve(prob, Tsit5(), reltol=1e-8, abstol=1e-8)
    assert sol.retcode == ReturnCode.Success
    assert len(sol.t) > 0 and len(sol.u) > 0
    assert sol.u[0][0] == 0.5
    assert abs(sol.u[-1][0] - 0.5 * math.exp(1.01)) < 1e-7


def test_core_saveat_and_lorenz():
    say("core.jl saveat + lorenz")
    prob = ODEProblem(mideq.exp_growth, 0.5, (0.0, 1.0), [1.01])
    sol = solve(prob, Tsit5(), saveat=0.1)
    assert sol.retcode == ReturnCode.Success and len(sol.t) == 11
    lp = ODEProblem(mideq.lorenz, [1.0, 0.0, 0.0], (0.0, 1.0), [10.0, 28.0, 8 / 3])
    sol = solve(lp, Tsit5())
    assert sol.retcode == ReturnCode.Success
    np.testing.assert_array_equal(sol.u[0], [1.0, 0.0, 0.0])
    assert len(sol.u[-1]) == 3
    rp = ODEProblem(mideq.robertson, [1.0, 0.0, 0.0], (0.0, 1e5), [0.04, 3e7, 1e4])
    assert solve(rp, Rodas5()).retcode == ReturnCode.Success
    assert solve(remake(prob, u0=1.0), Tsit5()).u[0][0] == 1.0


def test_ensemble_prob_func_output_func_rerun_reduction(oracle):
    say("ensemble prob_func/output_func/rerun/reduction")
    base = ODEProblem(mideq.lorenz, [1.0, 0.0, 0.0], (0.0, 2.0), [10.0, 28.0, 8 / 3])
    rng = np.random.default_rng(0)
    P = np.array([10.0, 28.0, 8 / 3])[:, None] * (1 + rng.uniform(-0.1, 0.1, (3, 40)))

    def prob_func(prob, i, repeat):
        return remake(prob, p=P[:, i - 1] * (1.0 if repeat == 1 else 1.01))

    seen = []

    def output_func(sol, i):
        seen.append(i)
        # rerun every 7th trajectory once
        return sol.u[-1], (i % 7 == 0 and sol.stats.naccept > 0 and len(seen) <= 40)

    def reduction(u, batch, I):
        return u + [float(np.sum(b)) for b in batch], False

    ens = EnsembleProblem(base, prob_func=prob_func, output_func=output_func,
                          reduction=reduction, u_init=[])
    es = solve(ens, Tsit5(), EnsembleMI355X(), trajectories=40, abstol=1e-9, reltol=1e-9)
    assert len(es.u) == 40
    assert sorted(seen[:40]) == list(range(1, 41)) and len(seen) == 40 + 5
    # rerun trajectories used repeat=2 parameters
    for i in range(1, 41):
        pi = P[:, i - 1] * (1.01 if i % 7 == 0 else 1.0)
        r = oracle.solve("lorenz", [1, 0, 0], pi, (0.0, 2.0), abstol=1e-9, reltol=1e-9)
        assert es.u[i - 1] == pytest.approx(float(np.sum(r["u"])), rel=1e-14)


def test_vectorized_prob_func_large():
    say("vectorized prob_func N=100000")
    w = mideq.workloads.lorenz(100000, seed=3)

    def pf(prob, I, repeat):
        return dict(u0=w["u0"], p=w["p"])
    pf.vectorized = True
    base = ODEProblem(mideq.lorenz, [1.0, 0.0, 0.0], (0.0, 10.0), [10.0, 28.0, 8 / 3])
    es = solve(EnsembleProblem(base, prob_func=pf, output_func=lambda s, i: (None, False)),
               Tsit5(), EnsembleMI355X(), trajectories=100000, abstol=1e-8, reltol=1e-8)
    a = es.arrays
    assert (a.retcode == 1).all()
    assert 500 < a.naccept.mean() < 1100


def test_c_abi_example_program():
    """differentialequations.jl_amd/examples/abi_example.c, built with the
    library: the C ABI from plain C (built-in model, run-time compiled model
    with a terminating callback and an on-device mean, a compile error)."""
    import os
    import subprocess
    say("C ABI example program")
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "differentialequations.jl_amd", "_build", "abi_example")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, (r.stdout, r.stderr)
    assert r.stdout.startswith("ok"), r.stdout


def test_ensemble_batch_size_and_converged():
    """solve(...; batch_size): trajectories in batches of batch_size, the
    reduction after each with its 1-based indices, stopping at the first
    batch it reports converged; the default reduction concatenates."""
    say("ensemble batch_size / converged")
    base = ODEProblem(mideq.lorenz, [1.0, 0.0, 0.0], (0.0, 1.0), [10.0, 28.0, 8 / 3])
    called = []

    def prob_func(prob, i, repeat):
        called.append(i)
        return remake(prob, u0=[1.0 + 0.01 * i, 0.0, 0.0])

    seen_I = []

    def reduction(u, batch, I):
        seen_I.append(list(I))
        u = u + [float(b.u[-1][0]) for b in batch]
        return u, len(u) >= 32

    ens = EnsembleProblem(base, prob_func=prob_func, reduction=reduction, u_init=[])
    es = solve(ens, Tsit5(), EnsembleMI355X(), trajectories=40, batch_size=16, abstol=1e-9,
               reltol=1e-9)
    assert seen_I == [list(range(1, 17)), list(range(17, 33))]
    assert sorted(called) == list(range(1, 33)) and es.converged is True and len(es.u) == 32
    one = solve(EnsembleProblem(base, prob_func=prob_func), Tsit5(), EnsembleMI355X(),
                trajectories=40, abstol=1e-9, reltol=1e-9)
    batched = solve(EnsembleProblem(base, prob_func=prob_func), Tsit5(), EnsembleMI355X(),
                    trajectories=40, batch_size=16, abstol=1e-9, reltol=1e-9)
    assert one.converged is True and batched.converged is False
    assert [s.u[-1][0] for s in one.u] == [s.u[-1][0] for s in batched.u]
    with pytest.raises(ValueError):
        solve(ens, Tsit5(), EnsembleMI355X(), trajectories=4, batch_size=8)


def test_single_problem_non_autonomous_backward(oracle):
    """The INTEGRATION.md example: one ODEProblem of a run-time compiled,
    non-autonomous model over tspan (10, 0) with saveat = 1.0: sol.t runs
    10, 9, ..., 0 (the order of integration), values bit-exact against the
    oracle."""
    import user_model_host as umh
    say("single problem, non-autonomous, backward, scalar saveat")
    rhs = "du[0] = u[1];\ndu[1] = -p[0] * u[0] - p[1] * u[1] + t * t;"
    forced = mideq.UserModel(2, 2, rhs)
    prob = ODEProblem(forced, [0.0, 0.0], (10.0, 0.0), [4.0, 0.1])
    sol = solve(prob, Tsit5(), saveat=1.0, abstol=1e-9, reltol=1e-9)
    assert sol.retcode == ReturnCode.Success
    np.testing.assert_array_equal(sol.t, np.arange(10.0, -1.0, -1.0))
    umh.install(oracle, 2, 2, rhs)
    c = oracle.solve("user", [0.0, 0.0], [4.0, 0.1], (10.0, 0.0), abstol=1e-9, reltol=1e-9,
                     saveat=np.arange(0.0, 11.0))
    np.testing.assert_array_equal(np.array(sol.u), c["saveat"][::-1])


def test_core_dense_output_interpolation(oracle):
    """core.jl:51-58: solve(prob, Tsit5(), dense = true); sol(0.5) > 0.5.
    sol(t) evaluates the interpolant of the step covering t: bit-exact
    against the same solve with t as saveat (on the device and in the
    oracle), for scalars, unsorted arrays with duplicates, NaN outside the
    span; solutions made with saveat have no dense output."""
    say("core.jl dense output sol(t)")
    prob = ODEProblem(mideq.exp_growth, 0.5, (0.0, 1.0), [1.01])
    sol = solve(prob, Tsit5(), dense=True)
    assert sol.retcode == ReturnCode.Success
    v = sol(0.5)
    assert v.shape == (1,) and v[0] > 0.5
    assert abs(v[0] - 0.5 * math.exp(0.505)) < 1e-3 * 0.5 * math.exp(0.505) + 1e-6
    ts = np.array([0.9, 0.1, 0.5, 0.5, 0.0, 1.0, 0.33])
    vals = sol(ts)
    assert vals.shape == (len(ts), 1)
    srt = np.sort(ts)
    ref = solve(prob, Tsit5(), saveat=srt)
    np.testing.assert_array_equal(vals[np.argsort(ts, kind="stable"), 0], np.array(ref.u)[:, 0])
    c = oracle.solve("exp", [0.5], [1.01], (0.0, 1.0), saveat=srt)
    np.testing.assert_array_equal(np.sort(vals[:, 0]), c["saveat"][:, 0])
    assert np.isnan(sol(np.array([-0.1, 1.5]))).all()
    # dense by default without saveat; off with saveat
    assert solve(prob, Tsit5())(0.25)[0] == solve(prob, Tsit5(), saveat=[0.25]).u[0][0]
    with pytest.raises(ValueError):
        solve(prob, Tsit5(), saveat=0.1)(0.5)


@pytest.mark.parametrize("model,alg", [("lorenz", Tsit5()), ("robertson", Rodas5()),
                                       ("brusselator", Rodas5())])
def test_ensemble_dense_output_matches_saveat(model, alg):
    """sim.u[i](t) of an ensemble equals the ensemble solved with t as
    saveat, trajectory by trajectory (lane kernels and the Brusselator team
    kernel, whose team shape the re-solve keeps)."""
    from mideq import workloads
    say(f"ensemble dense output {model}")
    N = 16
    if model == "brusselator":
        w = workloads.brusselator(N, seed=91)
        w["tf"] = np.minimum(w["tf"], 1.0)
        fn = mideq.brusselator
        base = ODEProblem(fn, w["u0"][:, 0][:int(w["nvec"][0])], (0.0, float(w["tf"][0])),
                          w["p"][:, 0])

        def pf(prob, i, repeat):
            n = int(w["nvec"][i - 1])
            return remake(prob, u0=w["u0"][:n, i - 1], p=w["p"][:, i - 1],
                          tspan=(0.0, float(w["tf"][i - 1])))
        ts = np.array([0.7, 0.05, 0.5])
    else:
        w = dict(lorenz=workloads.lorenz, robertson=workloads.robertson)[model](N, seed=92)
        fn = dict(lorenz=mideq.lorenz, robertson=mideq.robertson)[model]
        tf = 2.0 if model == "lorenz" else 10.0
        base = ODEProblem(fn, w["u0"][:, 0], (0.0, tf), w["p"][:, 0])

        def pf(prob, i, repeat):
            return remake(prob, u0=w["u0"][:, i - 1], p=w["p"][:, i - 1])
        ts = np.array([0.3, 1.7, 0.9]) * (1.0 if model == "lorenz" else 5.0)
    ens = EnsembleProblem(base, prob_func=pf)
    tol = dict(abstol=1e-8, reltol=1e-6)
    sim = solve(ens, alg, EnsembleMI355X(), trajectories=N, **tol)
    ref = solve(ens, alg, EnsembleMI355X(lanes=sim.arrays.lanes), trajectories=N,
                saveat=np.sort(ts), **tol)
    order = np.argsort(ts)
    for i in (0, 5, N - 1):
        vals = sim.u[i](ts)
        np.testing.assert_array_equal(vals[order], np.array(ref.u[i].u))


def test_dense_output_backward_span_and_terminate():
    """sol(t) of a solve integrated backward, and of one ended by a
    terminating callback: the values equal saveat solves at those times
    (bit for bit), NaN past the end of the solve."""
    from mideq import ContinuousCallback
    say("dense output, backward span and terminate")
    prob = ODEProblem(mideq.exp_growth, 2.0, (1.0, 0.0), [-1.5])
    sol = solve(prob, Tsit5(), abstol=1e-9, reltol=1e-9)
    ts = np.array([0.9, 0.1, 0.5, 1.0])
    ref = solve(prob, Tsit5(), saveat=np.sort(ts), abstol=1e-9, reltol=1e-9)
    # the backward solution lists its saved points in the order of integration
    np.testing.assert_array_equal(sol(np.sort(ts))[::-1, 0], np.array(ref.u)[:, 0])
    np.testing.assert_allclose(sol(0.5)[0], 2.0 * math.exp(-1.5 * (0.5 - 1.0)), rtol=1e-7)
    cb = ContinuousCallback("u[0] - 0.5", terminate=True)
    p2 = ODEProblem(mideq.exp_growth, 1.0, (0.0, 5.0), [-1.0])
    s2 = solve(p2, Tsit5(), callback=cb, abstol=1e-10, reltol=1e-10)
    assert s2.retcode == ReturnCode.Terminated
    tend = s2.t[-1]
    assert abs(tend - math.log(2.0)) < 1e-8
    v = s2(np.array([0.3, tend * 0.99, 1.0, 3.0]))
    np.testing.assert_allclose(v[:2, 0], np.exp(-np.array([0.3, tend * 0.99])), rtol=1e-8)
    assert np.isnan(v[2:]).all()


def test_core_save_everystep_single_problem(oracle):
    """solve(prob, Tsit5(); save_everystep=true) (upstream's default without
    saveat, core.jl:9-35): sol.t holds t0 and every accepted step's end,
    sol.u[1] == u0 and the last point is the final state, bit-exact against
    the oracle's step points."""
    import user_model_host as umh
    say("save_everystep single problem")
    prob = ODEProblem(mideq.UserModel(1, 1, "du[0] = p[0] * u[0];"), [0.5], (0.0, 1.0), [1.01])
    sol = solve(prob, Tsit5(), save_everystep=True, abstol=1e-8, reltol=1e-8)
    assert sol.retcode == ReturnCode.Success
    assert sol.t[0] == 0.0 and sol.t[-1] == 1.0 and sol.u[0][0] == 0.5
    assert len(sol.t) == sol.stats.naccept + 1 and (np.diff(sol.t) > 0).all()
    umh.install(oracle, 1, 1, "du[0] = p[0] * u[0];", save_everystep=True)
    r = oracle.ensemble("user", np.full((1, 1), 0.5), np.full((1, 1), 1.01), 0.0, 1.0,
                        abstol=1e-8, reltol=1e-8, event_cap=4096)
    m = int(r["ev_n"][0])
    np.testing.assert_array_equal(sol.t[1:], r["ev_t"][0, :m])
    np.testing.assert_array_equal(np.array([v[0] for v in sol.u[1:]]), r["ev_u"][0, :m, 0])


def test_save_start_and_save_end():
    """save_start / save_end (OrdinaryDiffEq's flags) on core.jl's saveat
    problem: saveat = 0.1 without the start keeps 10 points from 0.1; without
    saveat, save_start = false leaves the final point only and save_end =
    false the initial one."""
    say("save_start / save_end")
    f = mideq.UserModel(1, 1, "du[0] = p[0] * u[0];")
    prob = ODEProblem(f, [0.5], (0.0, 1.0), [1.01])
    full = solve(prob, Tsit5(), saveat=0.1, abstol=1e-8, reltol=1e-8)
    sol = solve(prob, Tsit5(), saveat=0.1, save_start=False, abstol=1e-8, reltol=1e-8)
    assert len(sol.t) == 10 and sol.t[0] == full.t[1] and sol.t[-1] == 1.0
    np.testing.assert_array_equal(np.array(sol.u), np.array(full.u[1:]))
    sol = solve(prob, Tsit5(), saveat=0.1, save_end=False, abstol=1e-8, reltol=1e-8)
    assert len(sol.t) == 10 and sol.t[-1] == full.t[-2]
    s1 = solve(prob, Tsit5(), save_start=False, abstol=1e-8, reltol=1e-8)
    s2 = solve(prob, Tsit5(), save_end=False, abstol=1e-8, reltol=1e-8)
    assert list(s1.t) == [1.0] and list(s2.t) == [0.0] and s2.u[0][0] == 0.5
    assert s1.u[0][0] == full.u[-1][0]


def test_dense_output_with_tstops():
    """sol(t) of a solve with tstops solves again with the same tstops, so
    it matches saveat values of that solve bit for bit."""
    say("dense output with tstops")
    f = mideq.UserModel(1, 1, "du[0] = -p[0] * u[0];")
    prob = ODEProblem(f, [1.0], (0.0, 1.0), [2.0])
    ts = [0.31, 0.62]
    sol = solve(prob, Tsit5(), tstops=ts, abstol=1e-9, reltol=1e-9)
    ref = solve(prob, Tsit5(), tstops=ts, saveat=[0.1, 0.45, 0.9], abstol=1e-9, reltol=1e-9)
    np.testing.assert_array_equal(np.array([sol(t)[0] for t in (0.1, 0.45, 0.9)]),
                                  np.array([u[0] for u in ref.u]))
