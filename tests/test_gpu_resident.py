"""Device-resident solves (the zero-copy EnsembleSolution view of
test/qa/qa.jl:52): torch tensors in GPU memory in, torch tensors out through
mideq_ensemble_solve_device, per-trajectory solutions as views.  Every
result is compared bit for bit with the host entry point's on the same
inputs."""
import numpy as np
import pytest

from gpu_util import ALGS, FUNCS, say
from mideq import (ContinuousCallback, DeviceOutputFunc, DeviceReduction, EnsembleMI355X,
                   EnsembleProblem, ODEProblem, Tsit5, UserModel, device_solve, lorenz, remake,
                   solve, workloads)
from test_user_model import BALL

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FIELDS = ("retcode", "naccept", "nreject", "nf", "nstiff", "dt_first", "u_final", "t_final")


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _same(h, d, extra=()):
    for k in FIELDS + tuple(extra):
        hv, dv = getattr(h, k), getattr(d, k)
        assert torch.is_tensor(dv) and dv.is_cuda, k
        np.testing.assert_array_equal(dv.cpu().numpy(), hv, err_msg=k)
    if h.us is not None:
        np.testing.assert_array_equal(d.us.cpu().numpy(), h.us)


def _lorenz(N, seed):
    w = workloads.lorenz(N, seed=seed)
    t0, tf = np.zeros(N), np.full(N, w["tf"])
    return w["u0"], w["p"], t0, tf


@pytest.mark.parametrize("layout", ["soa", "trajectory"])
def test_resident_lorenz_equals_host_entry(layout):
    N = 3000
    say(f"resident lorenz saveat layout={layout} N={N}")
    u0, p, t0, tf = _lorenz(N, 301)
    sa = workloads.saveat_uniform(0.0, 10.0, 21)
    ealg = EnsembleMI355X(controller="fast", save_layout=layout)
    kw = dict(saveat=sa, abstol=1e-8, reltol=1e-8)
    h = device_solve(FUNCS["lorenz"], u0, p, t0, tf, None, ALGS["tsit5"], ealg, **kw)
    d = device_solve(FUNCS["lorenz"], _dev(u0), _dev(p), _dev(t0), _dev(tf), None,
                     ALGS["tsit5"], ealg, **kw)
    torch.cuda.synchronize()
    assert d.resident and not h.resident
    _same(h, d)
    # solution(i): the same times, states that are views of the result's tensor
    base = d.us.untyped_storage().data_ptr()
    end = base + d.us.untyped_storage().nbytes()
    for i in (0, 17, N - 1):
        sh, sd = h.solution(i), d.solution(i)
        np.testing.assert_array_equal(sd.t, sh.t)
        assert len(sd.u) == len(sh.u)
        for vh, vd in zip(sh.u, sd.u):
            assert torch.is_tensor(vd) and vd.is_cuda
            assert base <= vd.data_ptr() < end
            np.testing.assert_array_equal(vd.cpu().numpy(), vh)
        assert sd.retcode == sh.retcode and sd.stats == sh.stats


def test_resident_side_stream_and_device_saveat():
    """Inputs made on the current stream, the solve enqueued on another
    stream (ordered after them), saveat given as a GPU tensor; no saveat
    points: solution(i) is (t0, u0) and (tf, u_final) as views."""
    N = 1000
    say(f"resident lorenz side stream N={N}")
    u0, p, t0, tf = _lorenz(N, 302)
    sa = workloads.saveat_uniform(0.0, 10.0, 11)
    ealg = EnsembleMI355X()
    kw = dict(abstol=1e-9, reltol=1e-9)
    h = device_solve(FUNCS["lorenz"], u0, p, t0, tf, None, ALGS["tsit5"], ealg, saveat=sa, **kw)
    s = torch.cuda.Stream()
    d = device_solve(FUNCS["lorenz"], _dev(u0), _dev(p), _dev(t0), _dev(tf), None,
                     ALGS["tsit5"], ealg, saveat=_dev(sa), stream=s, **kw)
    s.synchronize()
    _same(h, d)
    h2 = device_solve(FUNCS["lorenz"], u0, p, t0, tf, None, ALGS["tsit5"], ealg, **kw)
    d2 = device_solve(FUNCS["lorenz"], _dev(u0), _dev(p), _dev(t0), _dev(tf), None,
                      ALGS["tsit5"], ealg, stream=s, **kw)
    s.synchronize()
    _same(h2, d2)
    sol = d2.solution(5)
    np.testing.assert_array_equal(sol.t, [0.0, h2.t_final[5]])
    assert sol.u[1].data_ptr() == d2.u_final[:, 5].data_ptr()
    np.testing.assert_array_equal(sol.u[1].cpu().numpy(), h2.u_final[:, 5])


def test_resident_brusselator_ragged_rodas5():
    """The team kernel through the device entry: a ragged batch (nvec on the
    GPU), Rodas5 with saveat; per-trajectory views cut to each size."""
    N = 96
    say(f"resident brusselator rodas5 N={N}")
    w = workloads.brusselator(N, seed=303)
    t0 = np.zeros(N)
    tf = np.minimum(np.asarray(w["tf"], dtype=np.float64), 2.0)
    sa = workloads.saveat_uniform(0.0, 2.0, 5)
    ealg = EnsembleMI355X()
    kw = dict(saveat=sa, abstol=1e-8, reltol=1e-6)
    h = device_solve(FUNCS["brusselator"], w["u0"], w["p"], t0, tf, w["nvec"], ALGS["rodas5"],
                     ealg, **kw)
    d = device_solve(FUNCS["brusselator"], _dev(w["u0"]), _dev(w["p"]), _dev(t0), _dev(tf),
                     _dev(w["nvec"]), ALGS["rodas5"], ealg, **kw)
    torch.cuda.synchronize()
    assert (h.retcode == 1).all()
    _same(h, d)
    for i in (0, 50):
        sh, sd = h.solution(i), d.solution(i)
        np.testing.assert_array_equal(sd.t, sh.t)
        for vh, vd in zip(sh.u, sd.u):
            assert vd.shape == vh.shape == (int(w["nvec"][i]),)
            np.testing.assert_array_equal(vd.cpu().numpy(), vh)


def test_resident_callbacks_event_points_and_reduction():
    """A bouncing ball with points saved at each bounce, an on-device
    output_func and its reduction: every array (event times, states,
    counts, outputs, the folded values) bit-equal to the host entry's, and
    solution(i) merges the event points the same way."""
    N = 512
    say(f"resident bouncing ball events + output + reduction N={N}")
    cb = ContinuousCallback("u[0]", affect="u[1] = -p[1] * u[1];", direction=-1,
                            save_positions=(True, True))
    out = DeviceOutputFunc("out[0] = u[1];\nout[1] = t;", 2)
    model = UserModel(2, 2, BALL)
    rng = np.random.default_rng(304)
    u0 = np.vstack([rng.uniform(5.0, 20.0, N), np.zeros(N)])
    p = np.vstack([np.full(N, 9.81), rng.uniform(0.8, 0.95, N)])
    t0, tf = np.zeros(N), np.full(N, 6.0)
    sa = np.linspace(0.0, 6.0, 7)
    ealg = EnsembleMI355X()
    kw = dict(saveat=sa, abstol=1e-10, reltol=1e-10, callback=cb, output=out, reduction="sum",
              event_cap=8)
    h = device_solve(model, u0, p, t0, tf, None, ALGS["tsit5"], ealg, **kw)
    d = device_solve(model, _dev(u0), _dev(p), _dev(t0), _dev(tf), None, ALGS["tsit5"], ealg,
                     **kw)
    torch.cuda.synchronize()
    assert (h.nevent > 0).all()
    _same(h, d, extra=("nevent", "t_event", "ev_t", "ev_u", "ev_n", "out", "reduced"))
    for i in (0, 9, N - 1):
        sh, sd = h.solution(i), d.solution(i)
        np.testing.assert_array_equal(sd.t, sh.t)
        for vh, vd in zip(sh.u, sd.u):
            np.testing.assert_array_equal(vd.cpu().numpy(), vh)


def test_resident_high_level_solve():
    """solve(EnsembleProblem, Tsit5(), EnsembleMI355X(resident=True)): the
    same ODESolutions as the host form (times, codes, counts; states bit-equal
    but as GPU views), and the on-device output_func / reduction path with
    its folded value left on the GPU."""
    N = 64
    say(f"resident high-level solve N={N}")
    rng = np.random.default_rng(305)
    P = np.vstack([rng.uniform(9, 11, N), rng.uniform(25, 30, N), rng.uniform(2.4, 2.9, N)])
    prob = ODEProblem(lorenz, [1.0, 0.0, 0.0], (0.0, 3.0), [10.0, 28.0, 8 / 3])
    ens = EnsembleProblem(prob, prob_func=lambda pr, i, repeat: remake(pr, p=P[:, i - 1]))
    kw = dict(trajectories=N, saveat=0.5, abstol=1e-9, reltol=1e-9)
    h = solve(ens, Tsit5(), EnsembleMI355X(), **kw)
    d = solve(ens, Tsit5(), EnsembleMI355X(resident=True), **kw)
    assert d.arrays.resident and len(d.u) == len(h.u) == N
    for a, b in zip(h.u, d.u):
        np.testing.assert_array_equal(b.t, a.t)
        assert b.retcode == a.retcode and b.stats == a.stats
        assert all(torch.is_tensor(v) and v.is_cuda for v in b.u)
        np.testing.assert_array_equal(torch.stack(b.u).cpu().numpy(), np.stack(a.u))
    out = DeviceOutputFunc("out[0] = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];", 1)
    ens2 = EnsembleProblem(prob, prob_func=ens.prob_func, output_func=out,
                           reduction=DeviceReduction("sum"))
    h2 = solve(ens2, Tsit5(), EnsembleMI355X(), trajectories=N, abstol=1e-9, reltol=1e-9)
    d2 = solve(ens2, Tsit5(), EnsembleMI355X(resident=True), trajectories=N, abstol=1e-9,
               reltol=1e-9)
    assert torch.is_tensor(d2.u) and d2.u.is_cuda
    np.testing.assert_array_equal(d2.u.cpu().numpy(), h2.u)
