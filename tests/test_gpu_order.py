"""Caller-given launch order (mideq_options.order): it changes which
trajectories share a wavefront, never a trajectory's result.  Every solve
below is compared with the oracle (natural order) or with the unordered
GPU solve: identical retcodes and step counts, bit-equal states."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import ALGS, FUNCS, assert_identical, cpu, gpu, say
from mideq import EnsembleMI355X, _abi, device_solve, workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("schedule", ["lane", "persistent"])
def test_lorenz_saveat_with_order_bitexact(oracle, schedule):
    N = 3000
    say(f"lorenz order {schedule} N={N}")
    w = workloads.lorenz(N, seed=71)
    sa = workloads.saveat_uniform(0.0, 10.0, 25)
    order = np.random.default_rng(5).permutation(N)
    g = gpu("lorenz", w["u0"], w["p"], 0.0, 10.0, schedule=schedule, saveat=sa,
            abstol=1e-8, reltol=1e-8, order=order)
    c = cpu(oracle, "lorenz", w["u0"], w["p"], 0.0, 10.0, saveat=sa, abstol=1e-8, reltol=1e-8)
    assert_identical(g, c)


@pytest.mark.parametrize("alg,controller", [("auto", "reference"), ("auto", "fast"),
                                            ("auto5p", "fast"), ("tsit5", "fast")])
def test_dt0_and_hint_match_oracle(oracle, alg, controller):
    """Per-trajectory dt0 and initial-method hints (what the learned predictor
    feeds) against the oracle given the same arrays."""
    N = 600
    say(f"vdp {alg} {controller} with dt0 + hint N={N}")
    w = workloads.vdp(N, seed=76)
    w["p"] = np.minimum(w["p"], 300.0)
    tf = 2.0 * w["p"][0]
    rng = np.random.default_rng(8)
    dt0 = 10.0 ** rng.uniform(-6, -2, N)
    hint = (rng.uniform(size=N) > 0.5).astype(np.int8)
    kw = dict(abstol=1e-8, reltol=1e-6, dt0=dt0, alg_hint=hint)
    g = gpu("vdp", w["u0"], w["p"], 0.0, tf, alg=alg, controller=controller,
            order=np.random.default_rng(9).permutation(N), **kw)
    c = cpu(oracle, "vdp", w["u0"], w["p"], 0.0, tf, alg=alg,
            controller=1 if controller == "fast" else 0, **kw)
    assert (c["retcode"] == 1).all()
    assert_identical(g, c)
    np.testing.assert_array_equal(g.nstiff, c["nstiff"])


def test_auto_hint_dt0_order_bitexact():
    N = 2500
    say(f"vdp auto with dt0, hint, order N={N}")
    w = workloads.vdp(N, seed=72)
    rng = np.random.default_rng(6)
    dt0 = 10.0 ** rng.uniform(-6, -2, N)
    hint = (rng.uniform(size=N) > 0.5).astype(np.int8)
    tf = np.asarray(w["tf"], dtype=np.float64)
    kw = dict(abstol=1e-8, reltol=1e-6, dt0=dt0, alg_hint=hint, maxiters=1_000_000)
    a = gpu("vdp", w["u0"], w["p"], 0.0, tf, alg="auto", **kw)
    order = np.argsort(-(a.naccept + a.nreject), kind="stable")  # longest first
    b = gpu("vdp", w["u0"], w["p"], 0.0, tf, alg="auto", order=order, **kw)
    for k in ("retcode", "naccept", "nreject", "nf", "nstiff", "u_final", "t_final", "dt_first"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=k)


def test_bruss_with_caller_order(oracle):
    from test_gpu_bruss import batch
    N = 200
    say(f"brusselator caller order N={N}")
    w = batch(N, seed=73, tf_max=3.0)
    tol = dict(abstol=1e-8, reltol=1e-6)
    order = np.arange(N)[::-1].copy()
    g = gpu("brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"], lanes=64,
            schedule="persistent", order=order, **tol)
    c = cpu(oracle, "brusselator", w["u0"], w["p"], 0.0, w["tf"], nvec=w["nvec"],
            norm_lanes=64, **tol)
    assert_identical(g, c, nvec=w["nvec"])


def test_host_entry_rejects_non_permutation():
    say("order: non-permutation rejected")
    w = workloads.lorenz(100, seed=74)
    order = np.arange(100)
    order[10] = 11  # duplicate, 10 missing
    with pytest.raises(_abi.MideqError, match="permutation"):
        gpu("lorenz", w["u0"], w["p"], 0.0, 1.0, order=order)
    with pytest.raises(ValueError):
        gpu("lorenz", w["u0"], w["p"], 0.0, 1.0, order=np.arange(99))


def test_device_entry_skips_out_of_range_entries():
    import torch
    N = 1000
    say(f"order: device entry, out-of-range entry skipped N={N}")
    w = workloads.lorenz(N, seed=75)
    dev = torch.device("cuda:0")
    u0 = torch.from_numpy(w["u0"]).to(dev)
    p = torch.from_numpy(w["p"]).to(dev)
    uf = torch.zeros((3, N), dtype=torch.float64, device=dev)
    rc = torch.full((N,), 77, dtype=torch.int8, device=dev)
    na = torch.zeros(N, dtype=torch.int32, device=dev)
    order_np = np.random.default_rng(7).permutation(N)
    missing = int(order_np[10])
    order_np[10] = N + 100          # out of range: skipped, trajectory `missing` not run
    order = torch.from_numpy(order_np).to(dev)
    P = _abi.Problem()
    P.model, P.n, P.np, P.ntraj = _abi.MODEL["lorenz"], 3, 3, N
    P.u0, P.p, P.tspan0, P.tspan1 = u0.data_ptr(), p.data_ptr(), 0.0, 2.0
    O = _abi.Options()
    O.alg, O.precision, O.abstol, O.reltol = _abi.ALG["tsit5"], 0, 1e-8, 1e-8
    O.qsteady_default, O.schedule, O.order = 1, _abi.SCHEDULE["lane"], order.data_ptr()
    out = _abi.Outputs()
    out.u_final, out.retcode, out.naccept = uf.data_ptr(), rc.data_ptr(), na.data_ptr()
    ctx = _abi.context(0)
    ctx.solve_device(P, O, out, None)
    torch.cuda.synchronize()
    ref = gpu("lorenz", w["u0"], w["p"], 0.0, 2.0, abstol=1e-8, reltol=1e-8)
    rch = rc.cpu().numpy()
    assert rch[missing] == 77 and na.cpu().numpy()[missing] == 0
    keep = np.arange(N) != missing
    np.testing.assert_array_equal(rch[keep], ref.retcode[keep])
    np.testing.assert_array_equal(uf.cpu().numpy()[:, keep], ref.u_final[:, keep])


def test_saveat_size_limit_rejected_before_any_access():
    """n * ntraj >= 2^32 with saveat is refused by validation, which runs
    before any input is read (the buffers here hold one trajectory)."""
    say("saveat size limit: n*ntraj >= 2^32 rejected")
    ctx = _abi.context(0)
    u0 = np.zeros(3)
    p = np.ones(3)
    sa = np.array([0.5])
    P = _abi.Problem()
    P.model = _abi.MODEL["lorenz"]
    P.n = 3
    P.np = 3
    P.ntraj = 1 << 31
    P.u0 = u0.ctypes.data
    P.p = p.ctypes.data
    P.tspan0 = 0.0
    P.tspan1 = 1.0
    O = _abi.Options()
    O.alg = _abi.ALG["tsit5"]
    O.abstol = O.reltol = 1e-6
    O.saveat = sa.ctypes.data
    O.ns = 1
    out = _abi.Outputs()
    us = np.zeros(3)
    out.us = us.ctypes.data
    with pytest.raises(_abi.MideqError, match="2\\^32"):
        ctx.solve(P, O, out)


def test_empty_ensemble_both_entries():
    """ntraj = 0 is a successful no-op through both entry points (no
    launch, no output written), with and without saveat."""
    import torch
    say("empty ensemble, host and device entry")
    ctx = _abi.context(0)
    for ns in (0, 3):
        sa = np.array([0.0, 0.5, 1.0])
        P = _abi.Problem()
        P.model = _abi.MODEL["lorenz"]
        P.n = 3
        P.np = 3
        P.ntraj = 0
        P.tspan0 = 0.0
        P.tspan1 = 1.0
        O = _abi.Options()
        O.alg = _abi.ALG["tsit5"]
        O.abstol = O.reltol = 1e-6
        if ns:
            O.saveat = sa.ctypes.data
            O.ns = ns
        out = _abi.Outputs()
        ctx.solve(P, O, out)
        if ns:
            O.saveat = torch.from_numpy(sa).cuda().data_ptr()
        ctx.solve_device(P, O, out, None)
        torch.cuda.synchronize()
