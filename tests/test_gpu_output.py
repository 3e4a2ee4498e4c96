"""On-device output_func and reduction (EnsembleProblem output_func /
reduction, qa.jl:50): the per-trajectory values against the same
expression on the returned final states, and the folds against their
stated order restated in numpy (bit-exact), without and with the per-
trajectory arrays coming back."""
import numpy as np
import pytest

from gpu_util import ALGS, say
from mideq import (DeviceOutputFunc, DeviceReduction, EnsembleMI355X, EnsembleProblem,
                   ODEProblem, Tsit5, device_solve, lorenz, solve, workloads)

pytestmark = pytest.mark.gpu

OUT = DeviceOutputFunc("out[0] = u[0];\nout[1] = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];\n"
                       "out[2] = t;\nout[3] = (double)rc;", nout=4)


def fold(x, op):
    """The library's fold order: blocks of 256 by a halving tree (identity
    past the end), then block results in order; mean = sum / N."""
    ident = {"sum": 0.0, "mean": 0.0, "min": np.inf, "max": -np.inf}[op]
    comb = {"sum": np.add, "mean": np.add, "min": np.fmin, "max": np.fmax}[op]
    N = len(x)
    nb = (N + 255) // 256
    y = np.full(nb * 256, ident)
    y[:N] = x
    y = y.reshape(nb, 256)
    w = 128
    while w >= 1:
        y[:, :w] = comb(y[:, :w], y[:, w:2 * w])
        w //= 2
    acc = y[0, 0]
    for b in range(1, nb):
        acc = comb(acc, y[b, 0])
    return acc / N if op == "mean" else acc


def _run(N, reduction=None, keep_states=True, seed=121):
    w = workloads.lorenz(N, seed=seed)
    return device_solve(lorenz, w["u0"], w["p"], np.zeros(N), np.full(N, 2.0), None, Tsit5(),
                        EnsembleMI355X(), abstol=1e-8, reltol=1e-8, output=OUT,
                        reduction=reduction, keep_states=keep_states), w


@pytest.mark.parametrize("N", [1, 255, 257, 100003])
def test_output_values_and_folds(N):
    say(f"device output + folds N={N}")
    outs = {}
    for op in ("sum", "mean", "min", "max"):
        r, w = _run(N, reduction=op)
        u = r.u_final
        np.testing.assert_array_equal(r.out[0], u[0])
        np.testing.assert_array_equal(r.out[1], u[0] * u[0] + u[1] * u[1] + u[2] * u[2])
        np.testing.assert_array_equal(r.out[2], r.t_final)
        np.testing.assert_array_equal(r.out[3], r.retcode.astype(np.float64))
        for k in range(4):
            assert r.reduced[k] == fold(r.out[k], op), (op, k)
        outs[op] = r.reduced
    np.testing.assert_allclose(outs["mean"] * N, outs["sum"], rtol=1e-14)


def test_only_the_reduction_comes_back():
    N = 200000
    say(f"device reduction only N={N}")
    full, w = _run(N, reduction="mean")
    r, _ = _run(N, reduction="mean", keep_states=False)
    assert r.u_final is None and r.out is None
    np.testing.assert_array_equal(r.reduced, full.reduced)
    # through the SciMLBase surface: EnsembleProblem(output_func, reduction)
    prob = ODEProblem(lorenz, [1.0, 0.0, 0.0], (0.0, 2.0), [10.0, 28.0, 8.0 / 3.0])

    def pf(prob, I, repeat):  # I: 1-based trajectory indices
        return dict(u0=w["u0"][:, I - 1], p=w["p"][:, I - 1], t0=np.zeros(len(I)),
                    tf=np.full(len(I), 2.0))
    pf.vectorized = True
    ens = EnsembleProblem(prob, prob_func=pf, output_func=OUT,
                          reduction=DeviceReduction("mean"))
    sol = solve(ens, Tsit5(), EnsembleMI355X(), trajectories=N, abstol=1e-8, reltol=1e-8)
    np.testing.assert_array_equal(sol.u, full.reduced)
