"""Helpers for the GPU parity tests: run the same seeded ensemble through
libmideq.so (device) and through the oracle (CPU) and compare."""
import os
import sys

import numpy as np

import mideq
from mideq import EnsembleMI355X, device_solve
from mideq.scimlbase import AutoTsit5, Rodas5, Rodas5P, Tsit5

ALGS = dict(tsit5=Tsit5(), rodas5=Rodas5(), auto=AutoTsit5(), rodas5p=Rodas5P(),
            auto5p=AutoTsit5(Rodas5P()))
FUNCS = dict(exp=mideq.exp_growth, lorenz=mideq.lorenz, robertson=mideq.robertson,
             vdp=mideq.vanderpol, linear=mideq.linear3, brusselator=mideq.brusselator)
NTHREADS = min(16, os.cpu_count() or 1)


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def gpu(model, u0, p, t0, tf, alg="tsit5", precision="fp64", schedule="lane", saveat=None,
        save_dtype="float64", nvec=None, lanes=0, graph=False, controller="reference",
        auto_passes=0, iter_cap=None, save_layout="soa", **kw):
    N = u0.shape[1]
    t0 = np.broadcast_to(np.asarray(t0, dtype=np.float64), (N,)).copy()
    tf = np.broadcast_to(np.asarray(tf, dtype=np.float64), (N,)).copy()
    ealg = EnsembleMI355X(precision=precision, schedule=schedule, save_dtype=save_dtype,
                          lanes=lanes, graph=graph, controller=controller,
                          auto_passes=auto_passes, iter_cap=iter_cap, save_layout=save_layout)
    return device_solve(FUNCS[model], np.ascontiguousarray(u0), np.ascontiguousarray(p), t0, tf,
                        nvec, ALGS[alg], ealg, saveat=saveat, **kw)


def cpu(oracle, model, u0, p, t0, tf, alg="tsit5", precision="fp64", saveat=None, nvec=None,
        norm_lanes=64, **kw):
    return oracle.ensemble(model, u0, p, t0, tf, nthreads=NTHREADS, alg=alg,
                           precision=precision, saveat=saveat, nvec=nvec,
                           norm_lanes=norm_lanes, **kw)


def assert_identical(g, c, save=True, f32=False, nvec=None):
    """FP64 parity: same retcodes, same accept/reject/f-eval counts, bit-equal
    final states and saveat values."""
    np.testing.assert_array_equal(g.retcode, c["retcode"])
    np.testing.assert_array_equal(g.naccept, c["naccept"])
    np.testing.assert_array_equal(g.nreject, c["nreject"])
    np.testing.assert_array_equal(g.nf, c["nf"])
    if nvec is not None:
        mask = np.arange(g.u_final.shape[0])[:, None] < np.asarray(nvec)[None, :]
        np.testing.assert_array_equal(np.where(mask, g.u_final, 0.0), np.where(mask, c["u"], 0.0))
    else:
        np.testing.assert_array_equal(g.u_final, c["u"])
    np.testing.assert_array_equal(g.t_final, c["t"])
    np.testing.assert_array_equal(g.dt_first, c["dt_first"])
    if save and g.us is not None:
        ref = c["saveat"].astype(np.float32) if f32 else c["saveat"]
        np.testing.assert_array_equal(g.us, ref)
