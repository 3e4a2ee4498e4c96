"""mideq_ensemble_solve_device_noalloc against mideq_ensemble_solve_device
on the same device buffers: identical outputs, two calls in a row with the
second on a smaller ensemble (the fixed scratch keeps the first call's
contents, so the queue counters must be cleared by each call); refusals of
what the fixed region cannot hold, before any launch."""
import numpy as np
import pytest

from gpu_util import say
from mideq import _abi, workloads

pytestmark = pytest.mark.gpu


class Dev:
    """Device inputs and outputs of one ensemble and their ABI structs."""

    def __init__(self, torch, model, w, tf, alg="tsit5", schedule="persistent", auto_passes=0,
                 saveat=None, nvec=None, lanes=0, precision="fp64"):
        dev = torch.device("cuda:0")
        n, N = w["u0"].shape
        self.t = {}
        t = self.t
        t["u0"] = torch.from_numpy(np.ascontiguousarray(w["u0"])).to(dev)
        t["p"] = torch.from_numpy(np.ascontiguousarray(w["p"])).to(dev)
        t["t0"] = torch.zeros(N, dtype=torch.float64, device=dev)
        t["tf"] = torch.from_numpy(np.broadcast_to(np.asarray(tf, np.float64), (N,)).copy()).to(dev)
        self.P = _abi.Problem()
        P = self.P
        P.model, P.n, P.np, P.ntraj = _abi.MODEL[model], n, w["p"].shape[0], N
        P.u0, P.p, P.t0, P.tf = (t["u0"].data_ptr(), t["p"].data_ptr(), t["t0"].data_ptr(),
                                 t["tf"].data_ptr())
        P.tspan0, P.tspan1 = 0.0, float(np.max(tf))
        if nvec is not None:
            t["nvec"] = torch.from_numpy(np.ascontiguousarray(nvec, dtype=np.int32)).to(dev)
            P.nvec = t["nvec"].data_ptr()
        self.O = _abi.Options()
        O = self.O
        O.alg, O.precision = _abi.ALG[alg], _abi.PRECISION[precision]
        O.abstol, O.reltol, O.qsteady_default = 1e-8, 1e-6, 1
        O.schedule, O.auto_passes, O.lanes = _abi.SCHEDULE[schedule], auto_passes, lanes
        self.ns = 0
        if saveat is not None:
            t["saveat"] = torch.from_numpy(np.asarray(saveat, np.float64)).to(dev)
            O.saveat, O.ns = t["saveat"].data_ptr(), len(saveat)
            self.ns = len(saveat)
        self.shape = (n, N)
        self.dev = dev
        self.torch = torch

    def outputs(self):
        torch = self.torch
        n, N = self.shape
        o = dict(uf=torch.full((n, N), -1.0, dtype=torch.float64, device=self.dev),
                 rc=torch.zeros(N, dtype=torch.int8, device=self.dev),
                 na=torch.zeros(N, dtype=torch.int32, device=self.dev),
                 nr=torch.zeros(N, dtype=torch.int32, device=self.dev),
                 ns=torch.zeros(N, dtype=torch.int32, device=self.dev))
        if self.ns:
            o["us"] = torch.full((self.ns, n, N), -1.0, dtype=torch.float64, device=self.dev)
        out = _abi.Outputs()
        out.u_final, out.retcode = o["uf"].data_ptr(), o["rc"].data_ptr()
        out.naccept, out.nreject, out.nstiff = (o["na"].data_ptr(), o["nr"].data_ptr(),
                                                o["ns"].data_ptr())
        if self.ns:
            out.us = o["us"].data_ptr()
        return out, o


def compare(a, b):
    for k in a:
        np.testing.assert_array_equal(a[k].cpu().numpy(), b[k].cpu().numpy(), err_msg=k)


def both(ctx, d):
    """(device entry outputs, no-alloc entry outputs) of one ensemble."""
    torch = d.torch
    out1, o1 = d.outputs()
    ctx.solve_device(d.P, d.O, out1, None)
    out2, o2 = d.outputs()
    ctx.solve_device_noalloc(d.P, d.O, out2, None)
    torch.cuda.synchronize()
    assert (o1["rc"].cpu().numpy() == 1).all()
    return o1, o2


CASES = {
    # persistent lane kernel (queue counter) with saveat
    "lorenz_persistent_saveat": lambda N: ("lorenz", workloads.lorenz(N, seed=301), 10.0,
                                           dict(saveat=np.linspace(0.0, 10.0, 11))),
    # two-pass AutoSwitch: queue counter, handoff count and list
    "vdp_auto_two_pass": lambda N: ("vdp", workloads.vdp(N, seed=302), None,
                                    dict(alg="auto", auto_passes=2)),
    # team kernel queue, 64-lane teams
    "brusselator_team": lambda N: ("brusselator", workloads.brusselator(N, seed=303), None,
                                   dict(lanes=64, alg="auto")),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_noalloc_matches_device_entry_twice(case):
    import torch

    ctx = _abi.context(0)
    for N in ((20_000, 3_001) if case != "brusselator_team" else (600, 97)):
        model, w, tf, kw = CASES[case](N)
        if tf is None:
            tf = np.minimum(w["tf"], 40.0 if model == "vdp" else 1.0)
        say(f"no-alloc vs device entry: {case} N={N}")
        d = Dev(torch, model, w, tf, nvec=w.get("nvec"), **kw)
        o1, o2 = both(ctx, d)
        compare(o1, o2)


def test_noalloc_refusals():
    """Refused before any launch: a two-pass solve whose handoff list does
    not fit the fixed region (8 bytes per trajectory), and graph = 1; the
    outputs stay untouched."""
    import torch

    ctx = _abi.context(0)
    N = (_abi_scratch_bytes() - 256) // 8 + 1
    say(f"no-alloc refusals: two-pass handoff of N={N} and graph=1")
    w = workloads.vdp(N, seed=304)
    d = Dev(torch, "vdp", w, np.minimum(w["tf"], 1.0), alg="auto", auto_passes=2)
    out, o = d.outputs()
    with pytest.raises(_abi.MideqError, match="scratch"):
        ctx.solve_device_noalloc(d.P, d.O, out, None)
    d2 = Dev(torch, "lorenz", workloads.lorenz(64, seed=305), 1.0)
    d2.O.graph = 1
    out2, o2 = d2.outputs()
    with pytest.raises(_abi.MideqError, match="graph"):
        ctx.solve_device_noalloc(d2.P, d2.O, out2, None)
    torch.cuda.synchronize()
    assert (o["na"].cpu().numpy() == 0).all() and (o2["na"].cpu().numpy() == 0).all()


def _abi_scratch_bytes():
    import re

    with open(_abi.HEADER) as f:
        m = re.search(r"#define MIDEQ_NOALLOC_SCRATCH_BYTES \((\d+)u << (\d+)\)", f.read())
    return int(m.group(1)) << int(m.group(2))
