"""What the trained dt0 / method-class / cost network gains over the solver
without it: trained on the plain solver's logs of one seeded Lorenz +
Robertson + Van der Pol mix, evaluated on another.  The predicted class must
match the plain AutoTsit5(Rodas5()) verdict; starting each trajectory from
the predicted dt0 and method must cut rejected steps while every trajectory
still succeeds; and the launch order built from the predicted class and cost
must cut the lane slots that wavefronts waste waiting for their longest
lane (counted exactly from the per-trajectory step counts, so the check
does not depend on timing).  RHS evaluations and kernel times are printed
for the record."""
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


def wasted_fraction(cost, order=None):
    """Share of lane slots idle in 64-lane wavefronts that each run as long
    as their most expensive trajectory."""
    c = np.asarray(cost, dtype=np.float64)
    if order is not None:
        c = c[order]
    pad = (-len(c)) % 64
    c = np.concatenate([c, np.zeros(pad)]).reshape(-1, 64)
    slots = 64.0 * c.max(axis=1).sum()
    return 1.0 - c.sum() / slots


def solve(model, w, dt0=None, hint=None, order=None):
    import mideq
    from mideq import EnsembleMI355X, device_solve
    from mideq.scimlbase import AutoTsit5
    f = dict(lorenz=mideq.lorenz, robertson=mideq.robertson, vdp=mideq.vanderpol)[model]
    N = w["u0"].shape[1]
    at, rt = TOL[model]
    return device_solve(f, w["u0"], w["p"], np.zeros(N), np.broadcast_to(w["tf"], (N,)).copy(),
                        None, AutoTsit5(), EnsembleMI355X(), abstol=at, reltol=rt,
                        dt0=dt0, alg_hint=hint, order=order)


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
    Xs, ts, ls, cs = [], [], [], []
    for m, w in train.items():
        r = solve(m, w)
        assert (r.retcode == 1).all()
        t, l, c = training_targets(r.dt_first, r.naccept, r.nstiff, r.nreject)
        Xs.append(feats(m, w))
        ts.append(t)
        ls.append(l)
        cs.append(c)
    X = np.concatenate(Xs)
    T = np.concatenate(ts)
    Lb = np.concatenate(ls)
    Cb = np.concatenate(cs)
    net = MLP(seed=0)
    net.set_norm(X.mean(0), X.std(0) + 1e-6)
    Xd, Td, Ld, Cd = (torch.from_numpy(a).cuda() for a in (X, T, Lb, Cb))
    net.fit(Xd, Td, Ld, Cd, epochs=8, batch=2048, lr=3e-3)
    evals = mix(4000, seed=202)
    tot = dict(plain_rej=0, learned_rej=0, plain_nf=0, learned_nf=0, agree=0, n=0,
               waste_natural=0.0, waste_learned=0.0, slots=0.0)
    for m, w in evals.items():
        plain = solve(m, w)
        Yd = net.infer(torch.from_numpy(feats(m, w)).cuda())
        order = net.launch_order(Yd).cpu().numpy()
        Y = Yd.cpu().numpy()
        dt0 = 10.0 ** Y[:, 0].astype(np.float64)
        hint = (Y[:, 1] > 0).astype(np.int8)
        learned = solve(m, w, dt0=dt0, hint=hint, order=order)
        assert (learned.retcode == 1).all(), m
        # ordering never changes a trajectory's result
        same = solve(m, w, dt0=dt0, hint=hint)
        np.testing.assert_array_equal(same.u_final, learned.u_final)
        np.testing.assert_array_equal(same.naccept, learned.naccept)
        _, lab, cost_true = training_targets(plain.dt_first, plain.naccept, plain.nstiff,
                                             plain.nreject)
        work = 2.0 ** cost_true.astype(np.float64)
        wn, wl = wasted_fraction(work), wasted_fraction(work, order)
        tot["waste_natural"] += wn * work.sum()
        tot["waste_learned"] += wl * work.sum()
        tot["slots"] += work.sum()
        say(f"{m}: idle lane share natural {wn:.3f} -> learned order {wl:.3f}")
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
    # the learned order wastes at most 60 % of the idle lane slots of the natural order
    assert tot["waste_learned"] < 0.6 * tot["waste_natural"], tot
