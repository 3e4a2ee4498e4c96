"""What the trained dt0 / method-class network gains over the solver without
it: trained on the plain solver's logs of one seeded Lorenz + Robertson +
Van der Pol mix, evaluated on another.  The predicted class must match the
plain AutoTsit5(Rodas5()) verdict, and starting each trajectory from the
predicted dt0 and method must cut rejected steps while every trajectory
still succeeds.  RHS evaluations and kernel times are printed for the
record (the method hint only sets where AutoTsit5 starts, so RHS work stays
within a few percent either way on this mix)."""
import numpy as np
import pytest

from gpu_util import say

pytestmark = pytest.mark.gpu

TOL = dict(lorenz=(1e-8, 1e-8), robertson=(1e-8, 1e-6), vdp=(1e-8, 1e-6))


def mix(n_each, seed):
    from mideq import workloads
    out = {}
    out["lorenz"] = workloads.lorenz(n_each, seed=seed)
    out["robertson"] = workloads.robertson(n_each, seed=seed)
    v = workloads.vdp(n_each, seed=seed)
    v["p"] = np.minimum(v["p"], 1000.0)  # bound the test's run time
    v["tf"] = 2.0 * v["p"][0]
    out["vdp"] = v
    return out


def solve(model, w, dt0=None, hint=None):
    import mideq
    from mideq import EnsembleMI355X, device_solve
    from mideq.scimlbase import AutoTsit5
    f = dict(lorenz=mideq.lorenz, robertson=mideq.robertson, vdp=mideq.vanderpol)[model]
    N = w["u0"].shape[1]
    at, rt = TOL[model]
    return device_solve(f, w["u0"], w["p"], np.zeros(N), np.broadcast_to(w["tf"], (N,)).copy(),
                        None, AutoTsit5(), EnsembleMI355X(), abstol=at, reltol=rt,
                        dt0=dt0, alg_hint=hint)


def feats(model, w):
    from mideq.learned import features
    at, rt = TOL[model]
    N = w["u0"].shape[1]
    return features(model, w["p"], w["u0"], 0.0, w["tf"], at, rt)


def test_learned_dt0_and_method_beat_plain():
    import torch
    from mideq.learned import MLP, training_targets
    say("learned: plain logs -> train -> evaluate")
    train = mix(6000, seed=101)
    Xs, ts, ls = [], [], []
    for m, w in train.items():
        r = solve(m, w)
        assert (r.retcode == 1).all()
        t, l = training_targets(r.dt_first, r.naccept, r.nstiff)
        Xs.append(feats(m, w))
        ts.append(t)
        ls.append(l)
    X = np.concatenate(Xs)
    T = np.concatenate(ts)
    Lb = np.concatenate(ls)
    net = MLP(seed=0)
    net.set_norm(X.mean(0), X.std(0) + 1e-6)
    Xd, Td, Ld = (torch.from_numpy(a).cuda() for a in (X, T, Lb))
    net.fit(Xd, Td, Ld, epochs=8, batch=2048, lr=3e-3)
    evals = mix(4000, seed=202)
    tot = dict(plain_rej=0, learned_rej=0, plain_nf=0, learned_nf=0, agree=0, n=0)
    for m, w in evals.items():
        plain = solve(m, w)
        Y = net.infer(torch.from_numpy(feats(m, w)).cuda()).cpu().numpy()
        dt0 = 10.0 ** Y[:, 0].astype(np.float64)
        hint = (Y[:, 1] > 0).astype(np.int8)
        learned = solve(m, w, dt0=dt0, hint=hint)
        assert (learned.retcode == 1).all(), m
        _, lab = training_targets(plain.dt_first, plain.naccept, plain.nstiff)
        tot["agree"] += int(((lab > 0.5) == (hint == 1)).sum())
        tot["n"] += len(hint)
        tot["plain_rej"] += int(plain.nreject.sum())
        tot["learned_rej"] += int(learned.nreject.sum())
        tot["plain_nf"] += int(plain.nf.sum())
        tot["learned_nf"] += int(learned.nf.sum())
        say(f"{m}: rejects {plain.nreject.sum()} -> {learned.nreject.sum()}, "
            f"nf {plain.nf.sum()} -> {learned.nf.sum()}, "
            f"kernel ms {plain.kernel_ms:.2f} -> {learned.kernel_ms:.2f}")
    say(str(tot))
    assert tot["agree"] / tot["n"] > 0.95
    assert tot["learned_rej"] < tot["plain_rej"]
    assert tot["learned_nf"] < 1.05 * tot["plain_nf"]
