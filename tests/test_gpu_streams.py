"""Device-entry solves in flight on two streams of one context at the same
time: each stream's solves use their own work-queue counters, handoff list
and cached graph, so both give the results of solving alone."""
import numpy as np
import pytest

from gpu_util import gpu, say
from mideq import _abi, workloads

pytestmark = pytest.mark.gpu


def device_problem(torch, model, w, tf, alg, schedule, auto_passes=0, graph=False, nvec=None,
                   lanes=0):
    """Device buffers and the ABI structs for one ensemble; the tensors are
    returned with the structs so that they stay alive."""
    dev = torch.device("cuda:0")
    n, N = w["u0"].shape
    keep = {}
    keep["u0"] = torch.from_numpy(np.ascontiguousarray(w["u0"])).to(dev)
    keep["p"] = torch.from_numpy(np.ascontiguousarray(w["p"])).to(dev)
    keep["tf"] = torch.from_numpy(np.broadcast_to(np.asarray(tf, np.float64), (N,)).copy()).to(dev)
    keep["t0"] = torch.zeros(N, dtype=torch.float64, device=dev)
    keep["uf"] = torch.zeros((n, N), dtype=torch.float64, device=dev)
    keep["rc"] = torch.zeros(N, dtype=torch.int8, device=dev)
    keep["na"] = torch.zeros(N, dtype=torch.int32, device=dev)
    keep["nr"] = torch.zeros(N, dtype=torch.int32, device=dev)
    P = _abi.Problem()
    P.model, P.n, P.np, P.ntraj = _abi.MODEL[model], n, w["p"].shape[0], N
    P.u0, P.p = keep["u0"].data_ptr(), keep["p"].data_ptr()
    P.t0, P.tf = keep["t0"].data_ptr(), keep["tf"].data_ptr()
    P.tspan0, P.tspan1 = 0.0, float(np.max(tf))
    if nvec is not None:
        keep["nvec"] = torch.from_numpy(np.ascontiguousarray(nvec, dtype=np.int32)).to(dev)
        P.nvec = keep["nvec"].data_ptr()
    O = _abi.Options()
    O.alg, O.precision = _abi.ALG[alg], 0
    O.abstol, O.reltol = 1e-8, 1e-6
    O.qsteady_default = 1
    O.schedule = _abi.SCHEDULE[schedule]
    O.auto_passes = auto_passes
    O.graph = 1 if graph else 0
    O.lanes = lanes
    out = _abi.Outputs()
    out.u_final, out.retcode = keep["uf"].data_ptr(), keep["rc"].data_ptr()
    out.naccept, out.nreject = keep["na"].data_ptr(), keep["nr"].data_ptr()
    return P, O, out, keep


def check(keep, ref, nvec=None):
    uf = keep["uf"].cpu().numpy()
    if nvec is not None:
        mask = np.arange(uf.shape[0])[:, None] < np.asarray(nvec)[None, :]
        uf = np.where(mask, uf, 0.0)
        want = np.where(mask, ref.u_final, 0.0)
    else:
        want = ref.u_final
    np.testing.assert_array_equal(keep["rc"].cpu().numpy(), ref.retcode)
    np.testing.assert_array_equal(keep["na"].cpu().numpy(), ref.naccept)
    np.testing.assert_array_equal(keep["nr"].cpu().numpy(), ref.nreject)
    np.testing.assert_array_equal(uf, want)


@pytest.mark.parametrize("graph", [False, True])
def test_two_streams_in_flight(graph):
    """Persistent-queue Tsit5 on one stream and two-pass AutoSwitch (queue
    counter, handoff count and list) on another, enqueued back to back
    before either is waited for, three rounds in a row."""
    import torch

    say(f"two streams of one context in flight, graph={graph}")
    ctx = _abi.context(0)
    wa = workloads.lorenz(40_000, seed=91)
    wb = workloads.vdp(30_000, seed=92)
    tfb = np.minimum(wb["tf"], 60.0)
    Pa, Oa, outa, ka = device_problem(torch, "lorenz", wa, 10.0, "tsit5", "persistent",
                                      graph=graph)
    Pb, Ob, outb, kb = device_problem(torch, "vdp", wb, tfb, "auto", "persistent",
                                      auto_passes=2, graph=graph)
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    ra = gpu("lorenz", wa["u0"], wa["p"], 0.0, 10.0, schedule="persistent",
             abstol=1e-8, reltol=1e-6)
    rb = gpu("vdp", wb["u0"], wb["p"], 0.0, tfb, alg="auto", schedule="persistent",
             auto_passes=2, abstol=1e-8, reltol=1e-6)
    for rnd in range(3):
        for k in (ka, kb):
            k["uf"].zero_()
            k["na"].zero_()
        torch.cuda.synchronize()
        ctx.solve_device(Pa, Oa, outa, sa.cuda_stream)
        ctx.solve_device(Pb, Ob, outb, sb.cuda_stream)
        torch.cuda.synchronize()
        check(ka, ra)
        check(kb, rb)


def test_two_streams_brusselator_and_lane():
    """The team kernel's queue on one stream beside a persistent lane solve
    on another."""
    import torch

    say("brusselator + lorenz on two streams of one context")
    ctx = _abi.context(0)
    wa = workloads.brusselator(2_000, seed=93)
    tfa = np.minimum(wa["tf"], 2.0)
    wb = workloads.lorenz(50_000, seed=94)
    Pa, Oa, outa, ka = device_problem(torch, "brusselator", wa, tfa, "tsit5", "persistent",
                                      nvec=wa["nvec"], lanes=64)
    Pb, Ob, outb, kb = device_problem(torch, "lorenz", wb, 10.0, "tsit5", "persistent")
    ra = gpu("brusselator", wa["u0"], wa["p"], 0.0, tfa, nvec=wa["nvec"], lanes=64,
             schedule="persistent", abstol=1e-8, reltol=1e-6)
    rb = gpu("lorenz", wb["u0"], wb["p"], 0.0, 10.0, schedule="persistent",
             abstol=1e-8, reltol=1e-6)
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    for rnd in range(2):
        ctx.solve_device(Pa, Oa, outa, sa.cuda_stream)
        ctx.solve_device(Pb, Ob, outb, sb.cuda_stream)
        torch.cuda.synchronize()
        check(ka, ra, nvec=wa["nvec"])
        check(kb, rb)


@pytest.mark.parametrize("save_dtype", ["float64", "float32"])
def test_host_entry_large_output_into_pageable_memory(save_dtype):
    """Saved values larger than the staging threshold reach pageable caller
    memory through the pinned staging workers of the host entry point: the
    same bytes as a copy into a pinned destination (torch's pinned
    allocation) of the same solve."""
    import torch
    N, ns = 131072, 64
    say(f"host entry, {save_dtype} saveat output of {N}x{ns}x3 into pageable memory")
    w = workloads.lorenz(N, seed=121)
    sa = np.linspace(0.0, 2.0, ns)
    ctx = _abi.context(0)

    def solve(us):
        P = _abi.Problem()
        P.model, P.n, P.np, P.ntraj = _abi.MODEL["lorenz"], 3, 3, N
        u0 = np.ascontiguousarray(w["u0"])
        p = np.ascontiguousarray(w["p"])
        P.u0, P.p, P.tspan0, P.tspan1 = u0.ctypes.data, p.ctypes.data, 0.0, 2.0
        O = _abi.Options()
        O.alg, O.precision, O.abstol, O.reltol = _abi.ALG["tsit5"], 0, 1e-8, 1e-8
        O.qsteady_default = 1
        O.saveat, O.ns = sa.ctypes.data, ns
        O.save_dtype = 1 if save_dtype == "float32" else 0
        out = _abi.Outputs()
        out.us = us.ctypes.data if isinstance(us, np.ndarray) else us.data_ptr()
        ctx.solve(P, O, out)

    dt = np.float32 if save_dtype == "float32" else np.float64
    pageable = np.full((ns, 3, N), np.nan, dtype=dt)
    assert pageable.nbytes > (64 << 20)
    solve(pageable)
    pinned = torch.empty((ns, 3, N), dtype=torch.float32 if dt == np.float32 else torch.float64,
                         pin_memory=True)
    solve(pinned)
    np.testing.assert_array_equal(pageable, pinned.numpy())
    assert not np.isnan(pageable).any()
