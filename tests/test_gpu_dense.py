"""saveat inside Rosenbrock steps (Rodas5, Rodas5P, AutoTsit5 in its stiff
phase): the lane kernels' dense output against the oracle's, bit-exact, in
both precisions, both controllers, both schedules and FP32 storage."""
import numpy as np
import pytest

from gpu_util import assert_identical, cpu, gpu, say
from mideq import workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("alg,precision,controller,schedule,save_dtype", [
    ("rodas5", "fp64", "reference", "lane", "float64"),
    ("rodas5", "fp64", "fast", "persistent", "float64"),
    ("rodas5p", "fp64", "reference", "lane", "float64"),
    ("rodas5p", "mixed", "fast", "lane", "float64"),
    ("rodas5", "fp64", "reference", "lane", "float32"),
])
def test_robertson_saveat_rosenbrock_bitexact(oracle, alg, precision, controller, schedule,
                                              save_dtype):
    N = 2000
    say(f"robertson saveat {alg} {precision} {controller} {schedule} {save_dtype} N={N}")
    w = workloads.robertson(N, seed=91)
    sa = np.concatenate([[0.0], np.logspace(-6, 5, 45)])
    kw = dict(abstol=1e-8, reltol=1e-6, saveat=sa)
    g = gpu("robertson", w["u0"], w["p"], 0.0, 1e5, alg=alg, precision=precision,
            controller=controller, schedule=schedule, save_dtype=save_dtype, **kw)
    c = cpu(oracle, "robertson", w["u0"], w["p"], 0.0, 1e5, alg=alg, precision=precision,
            controller=1 if controller == "fast" else 0, **kw)
    assert (c["retcode"] == 1).all()
    assert_identical(g, c, f32=save_dtype == "float32")
    # sum y = 1 is linear, so the FP64 interpolant keeps it too (FP32 stages
    # round each increment, FP32 storage each value)
    s = g.us.astype(np.float64).sum(axis=1)
    tol = 1e-4 if precision == "mixed" else 1e-6 if save_dtype == "float32" else 1e-12
    assert np.abs(s - 1.0).max() < tol


@pytest.mark.parametrize("alg", ["auto", "auto5p"])
def test_vdp_saveat_autoswitch_bitexact(oracle, alg):
    """Saved points fall in Tsit5 steps and in Rosenbrock steps of one solve."""
    N = 1000
    say(f"vdp saveat {alg} N={N}")
    w = workloads.vdp(N, seed=92)
    w["p"] = np.minimum(w["p"], 300.0)
    sa = np.linspace(0.0, 50.0, 26)
    kw = dict(abstol=1e-8, reltol=1e-6, saveat=sa)
    g = gpu("vdp", w["u0"], w["p"], 0.0, 50.0, alg=alg, **kw)
    c = cpu(oracle, "vdp", w["u0"], w["p"], 0.0, 50.0, alg=alg, **kw)
    assert (c["retcode"] == 1).all()
    assert ((c["nstiff"] > 0) & (c["nstiff"] < c["naccept"])).any()
    assert_identical(g, c)


def test_lorenz_saveat_autoswitch_tsit5_phase(oracle):
    """AutoTsit5 that never switches: the dense build's Tsit5 interpolation."""
    N = 1500
    say(f"lorenz saveat auto N={N}")
    w = workloads.lorenz(N, seed=93)
    sa = workloads.saveat_uniform(0.0, 10.0, 21)
    kw = dict(abstol=1e-8, reltol=1e-8, saveat=sa)
    g = gpu("lorenz", w["u0"], w["p"], 0.0, 10.0, alg="auto", **kw)
    c = cpu(oracle, "lorenz", w["u0"], w["p"], 0.0, 10.0, alg="auto", **kw)
    assert_identical(g, c)
    t = gpu("lorenz", w["u0"], w["p"], 0.0, 10.0, alg="tsit5", **kw)
    np.testing.assert_array_equal(g.us, t.us)
